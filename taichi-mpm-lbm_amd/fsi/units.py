"""Unit conversion placeholder, import-compatible with the reference's
``fsi/units.py:4-10``: every quantity in this package is in lattice units
(dx = dt = 1 on the lattice), so the class carries no conversion yet."""

from __future__ import annotations


class LatticeUnits:
    """Physical-to-lattice unit conversion (none: inputs are lattice units)."""
