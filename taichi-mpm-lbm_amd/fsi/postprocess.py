"""Readers for the NPZ snapshots (reference: fsi/postprocess.py:32-115).
Host-side numpy only; the reference's CSV/JSON writers, validation tables and
plotting are offline tooling around the step and are not part of this package.

Snapshots are `snapshot_<step>.npz` files written by fsi.output. The time
series derived from them use the reference's keys and definitions
(fsi/postprocess.py:241-289), e.g. `lbm_total_mass_estimate` = sum of the
density field and `mpm_center_of_mass_*` = the unweighted mean particle
position, so tables and plots built on either implementation agree.
"""

from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path
from typing import Any

import numpy as np

# derived time-series keys, in the reference's column order
_ORDER = (
    "step", "time", "mpm_particle_count", "lbm_mean_density", "lbm_total_mass_estimate",
    "lbm_max_velocity_norm", "mpm_center_of_mass_x", "mpm_center_of_mass_y", "mpm_center_of_mass_z",
    "mpm_max_velocity_norm", "coupling_force_norm", "total_particle_coupling_force_norm",
    "total_fluid_coupling_force_norm", "coupling_unsupported_particle_count",
    "coupling_partial_support_particle_count", "coupling_clipped_particle_count", "ib_force_norm",
    "ib_total_force_norm", "ib_active_cell_count", "ib_clipped_cell_count", "contact_candidate_count",
    "contact_damped_particle_count",
)
_COUNTS = ("coupling_unsupported_particle_count", "coupling_partial_support_particle_count",
           "coupling_clipped_particle_count", "ib_active_cell_count", "ib_clipped_cell_count",
           "contact_candidate_count", "contact_damped_particle_count")
# snapshot field -> derived key of its Frobenius norm
_NORMS = (("coupling_force", "coupling_force_norm"), ("immersed_boundary_force", "ib_force_norm"),
          ("total_particle_coupling_force", "total_particle_coupling_force_norm"),
          ("total_fluid_coupling_force", "total_fluid_coupling_force_norm"),
          ("ib_total_force", "ib_total_force_norm"))

@dataclass(frozen=True)
class SnapshotInfo:
    path: Path
    step: int
    time: float
    particle_count: int
    lbm_shape: tuple[int, int, int]

    def to_dict(self) -> dict[str, Any]:
        return {"path": str(self.path), "step": self.step, "time": self.time,
                "particle_count": self.particle_count, "lbm_shape": self.lbm_shape}


def _scalar(v, kind):
    return kind(np.asarray(v).reshape(()).item())


def _step_of(path: Path):
    tail = path.stem[len("snapshot_"):] if path.stem.startswith("snapshot_") else ""
    return (0, int(tail)) if tail.isdigit() else (1, path.name)


def list_npz_snapshots(output_dir: str | Path) -> list[Path]:
    """snapshot_*.npz of a directory, ordered by step number."""
    d = Path(output_dir)
    return sorted(d.glob("snapshot_*.npz"), key=_step_of) if d.exists() else []


def load_npz_snapshot(path: str | Path) -> dict[str, np.ndarray]:
    with np.load(Path(path), allow_pickle=False) as data:
        return {k: np.array(data[k]) for k in data.files}


def snapshot_info(path: str | Path) -> SnapshotInfo:
    data = load_npz_snapshot(path)
    shape = tuple(int(v) for v in np.asarray(data["lbm_shape"]).reshape(-1))
    if len(shape) != 3:
        raise ValueError(f"lbm_shape must have 3 entries, got {shape}.")
    return SnapshotInfo(path=Path(path), step=_scalar(data["step"], int), time=_scalar(data["time"], float),
                        particle_count=_scalar(data["mpm_particle_count"], int), lbm_shape=shape)


def summarize_snapshots(output_dir: str | Path) -> list[SnapshotInfo]:
    return [snapshot_info(p) for p in list_npz_snapshots(output_dir)]


def _max_norm(v) -> float:
    a = np.asarray(v, dtype=np.float64)
    return 0.0 if a.size == 0 else float(np.linalg.norm(a.reshape(-1, a.shape[-1]), axis=1).max())


def _row(path: Path) -> dict[str, float | int]:
    d = load_npz_snapshot(path)
    row: dict[str, float | int] = {"step": _scalar(d["step"], int), "time": _scalar(d["time"], float),
                                   "mpm_particle_count": _scalar(d["mpm_particle_count"], int)}
    if "lbm_density" in d:
        row["lbm_mean_density"] = float(np.mean(d["lbm_density"]))
        row["lbm_total_mass_estimate"] = float(np.sum(d["lbm_density"]))
    if "lbm_velocity" in d:
        row["lbm_max_velocity_norm"] = _max_norm(d["lbm_velocity"])
    if "mpm_positions" in d and np.asarray(d["mpm_positions"]).size:
        c = np.mean(np.asarray(d["mpm_positions"], dtype=np.float64), axis=0)
        row.update(mpm_center_of_mass_x=float(c[0]), mpm_center_of_mass_y=float(c[1]),
                   mpm_center_of_mass_z=float(c[2]))
    if "mpm_velocities" in d:
        row["mpm_max_velocity_norm"] = _max_norm(d["mpm_velocities"])
    for field, key in _NORMS:
        if field in d:
            row[key] = float(np.linalg.norm(np.asarray(d[field])))
    for key in _COUNTS:
        if key in d:
            row[key] = _scalar(d[key], int)
    return row


def _ordered(keys) -> list[str]:
    keys = set(keys)
    return [k for k in _ORDER if k in keys] + sorted(keys.difference(_ORDER))


def extract_snapshot_timeseries(output_dir: str | Path) -> dict[str, np.ndarray]:
    """One array per derived key present in every snapshot, in step order."""
    rows = [_row(p) for p in list_npz_snapshots(output_dir)]
    if not rows:
        return {"step": np.array([], dtype=np.int64), "time": np.array([], dtype=np.float64),
                "mpm_particle_count": np.array([], dtype=np.int64)}
    common = set.intersection(*(set(r) for r in rows))
    return {k: np.array([r[k] for r in rows]) for k in _ordered(common)}


def summarize_coupling_diagnostics(output_dir: str | Path) -> dict[str, np.ndarray]:
    ts = extract_snapshot_timeseries(output_dir)
    keys = ("step", "time", "coupling_force_norm", "total_particle_coupling_force_norm",
            "total_fluid_coupling_force_norm", *_COUNTS[:3], "ib_force_norm", "ib_total_force_norm",
            *_COUNTS[3:])
    return {k: ts[k] for k in keys if k in ts}
