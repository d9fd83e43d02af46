"""Snapshot writer (reference: fsi/output.py:15-198).

Both formats are driven by field tables: one row per array, naming the
output group that switches it on (``OutputConfig.write_lbm_fields``,
``write_mpm_particles``, ``write_coupling_fields``), the key under which it
is written, and how it is read from the simulation. The NPZ keys, dtypes and
shapes and the VTK file names, array names and grid spacing are the
reference's, so readers of its snapshots (fsi/postprocess.py) keep working.

VTK needs the optional ``pyevtk`` package, as in the reference. It is
imported when a VTK snapshot is written, so the NPZ path works without it.
"""

from __future__ import annotations

from pathlib import Path
from typing import TYPE_CHECKING, Callable

import numpy as np

from .config import OutputConfig

if TYPE_CHECKING:
    from .simulation import FSISimulation

_LBM, _MPM, _COUPLING = "write_lbm_fields", "write_mpm_particles", "write_coupling_fields"


def _counter(key: str) -> Callable:
    """An integer of coupling_diagnostics() as a 0-d int32 array."""
    return lambda sim, diag: np.array(diag[key], dtype=np.int32)


# NPZ snapshot, fsi/output.py:39-116: (group or None for always, key, reader).
# Readers take the simulation and its coupling_diagnostics() (computed once).
_NPZ_FIELDS: tuple[tuple[str | None, str, Callable], ...] = (
    (None, "step", lambda sim, diag: np.array(sim.step_index, dtype=np.int64)),
    (None, "time", lambda sim, diag: np.array(sim.time, dtype=np.float64)),
    (None, "lbm_shape", lambda sim, diag: np.array([sim.lbm.nx, sim.lbm.ny, sim.lbm.nz], dtype=np.int32)),
    (None, "mpm_particle_count", lambda sim, diag: np.array(sim.mpm.particle_count(), dtype=np.int32)),
    (_LBM, "lbm_density", lambda sim, diag: sim.lbm.density_numpy()),
    (_LBM, "lbm_velocity", lambda sim, diag: sim.lbm.velocity_numpy()),
    (_LBM, "lbm_force", lambda sim, diag: sim.lbm.force_numpy()),
    (_LBM, "lbm_solid", lambda sim, diag: sim.lbm.solid_numpy()),
    (_MPM, "mpm_positions", lambda sim, diag: sim.mpm.positions_numpy()),
    (_MPM, "mpm_velocities", lambda sim, diag: sim.mpm.velocities_numpy()),
    (_MPM, "mpm_particle_forces", lambda sim, diag: sim.mpm.particle_forces_numpy()),
    (_MPM, "mpm_deformation_gradients", lambda sim, diag: sim.mpm.deformation_gradients_numpy()),
    (_MPM, "mpm_active", lambda sim, diag: sim.mpm.active_numpy()),
    (_COUPLING, "coupling_force", lambda sim, diag: sim.coupler.coupling_force_numpy()),
    (_COUPLING, "solid_volume_fraction", lambda sim, diag: sim.coupler.solid_volume_fraction_numpy()),
    (_COUPLING, "immersed_boundary_force", lambda sim, diag: sim.coupler.immersed_boundary_force_numpy()),
    (_COUPLING, "total_particle_coupling_force", lambda sim, diag: sim.coupler.total_particle_coupling_force()),
    (_COUPLING, "total_fluid_coupling_force", lambda sim, diag: sim.coupler.total_fluid_coupling_force()),
    (_COUPLING, "coupling_particle_valid_weight", lambda sim, diag: sim.coupler.particle_valid_weights_numpy()),
    (_COUPLING, "coupling_particle_mask", lambda sim, diag: sim.coupler.particle_coupling_mask_numpy()),
    (_COUPLING, "particle_contact_mask", lambda sim, diag: sim.coupler.particle_contact_mask_numpy()),
    (_COUPLING, "ib_total_force",
     lambda sim, diag: np.array([diag[f"ib_total_force_{a}"] for a in "xyz"], dtype=np.float32)),
    (_COUPLING, "ib_active_cell_count", _counter("ib_active_cell_count")),
    (_COUPLING, "ib_clipped_cell_count", _counter("ib_clipped_cell_count")),
    (_COUPLING, "contact_candidate_count", _counter("contact_candidate_count")),
    (_COUPLING, "contact_damped_particle_count", _counter("contact_damped_particle_count")),
    (_COUPLING, "coupling_unsupported_particle_count", _counter("unsupported_particle_count")),
    (_COUPLING, "coupling_partial_support_particle_count", _counter("partial_support_particle_count")),
    (_COUPLING, "coupling_clipped_particle_count", _counter("clipped_particle_count")),
)

# VTK cell data of fluid_%06d.vti, fsi/output.py:128-166: (group, name, reader,
# vector). Vectors are written as their three component arrays.
_VTK_CELL_FIELDS: tuple[tuple[str, str, Callable, bool], ...] = (
    (_LBM, "rho", lambda sim: sim.lbm.density_numpy(), False),
    (_LBM, "velocity", lambda sim: sim.lbm.velocity_numpy(), True),
    (_LBM, "force", lambda sim: sim.lbm.force_numpy(), True),
    (_LBM, "solid", lambda sim: sim.lbm.solid_numpy(), False),
    (_COUPLING, "coupling_force", lambda sim: sim.coupler.coupling_force_numpy(), True),
    (_COUPLING, "solid_volume_fraction", lambda sim: sim.coupler.solid_volume_fraction_numpy(), False),
    (_COUPLING, "immersed_boundary_force", lambda sim: sim.coupler.immersed_boundary_force_numpy(), True),
)

# VTK point data of particles_%06d.vtu, fsi/output.py:168-189
_VTK_POINT_FIELDS: tuple[tuple[str | None, str, Callable, bool], ...] = (
    (None, "velocity", lambda sim: sim.mpm.velocities_numpy(), True),
    (None, "particle_force", lambda sim: sim.mpm.particle_forces_numpy(), True),
    (_COUPLING, "contact_mask", lambda sim: sim.coupler.particle_contact_mask_numpy(), False),
)


def _vtk_array(values: np.ndarray, vector: bool):
    if vector:
        return tuple(np.ascontiguousarray(values[..., a]) for a in range(3))
    return np.ascontiguousarray(values)


class SimulationOutputWriter:
    """Write lightweight simulation snapshots for analysis and visualization."""

    def __init__(self, config: OutputConfig):
        self.config = config
        self.config.validate()
        self.output_dir = Path(config.output_dir)

    def ensure_output_dir(self) -> None:
        self.output_dir.mkdir(parents=True, exist_ok=True)

    def should_write(self, step: int) -> bool:
        return int(step) % self.config.output_interval == 0

    def _enabled(self, group: str | None) -> bool:
        return group is None or bool(getattr(self.config, group))

    def write_snapshot(self, sim: "FSISimulation") -> list[Path]:
        writers = {
            "npz": lambda: [self.write_npz_snapshot(sim)],
            "vtk": lambda: self.write_vtk_snapshot(sim),
            "both": lambda: [self.write_npz_snapshot(sim), *self.write_vtk_snapshot(sim)],
        }
        if self.config.output_format not in writers:
            raise ValueError(f"Unsupported output format: {self.config.output_format}.")
        return writers[self.config.output_format]()

    def write_npz_snapshot(self, sim: "FSISimulation") -> Path:
        self.ensure_output_dir()
        path = self.output_dir / f"snapshot_{sim.step_index:06d}.npz"
        diag = sim.coupler.coupling_diagnostics() if self.config.write_coupling_fields else None
        arrays = {key: read(sim, diag) for group, key, read in _NPZ_FIELDS if self._enabled(group)}
        np.savez_compressed(path, **arrays)
        return path

    def write_vtk_snapshot(self, sim: "FSISimulation") -> list[Path]:
        from pyevtk.hl import imageToVTK, pointsToVTK  # ImportError when pyevtk is absent

        self.ensure_output_dir()
        step = sim.step_index
        paths: list[Path] = []
        if self.config.write_lbm_fields or self.config.write_coupling_fields:
            cells = {name: _vtk_array(read(sim), vec) for group, name, read, vec in _VTK_CELL_FIELDS
                     if self._enabled(group)}
            dx = float(sim.mpm.dx)
            paths.append(Path(imageToVTK(str(self.output_dir / f"fluid_{step:06d}"), origin=(0.0, 0.0, 0.0),
                                         spacing=(dx, dx, dx), cellData=cells)))
        if self.config.write_mpm_particles:
            pos = sim.mpm.positions_numpy()
            xyz = [np.ascontiguousarray(pos[:, a]) if len(pos) else np.zeros(0, np.float32) for a in range(3)]
            points = {name: _vtk_array(read(sim), vec) for group, name, read, vec in _VTK_POINT_FIELDS
                      if self._enabled(group)}
            paths.append(Path(pointsToVTK(str(self.output_dir / f"particles_{step:06d}"), *xyz, data=points)))
        return paths
