"""Output-folder helper, import-compatible with the reference's
``fsi/io.py:6-7`` (``ensure_output_dir``)."""

from __future__ import annotations

from pathlib import Path


def ensure_output_dir(path: Path) -> None:
    """Create ``path`` and any missing parents; an existing folder is kept."""
    Path(path).mkdir(parents=True, exist_ok=True)
