"""Snapshot writer (reference: fsi/output.py:15-198).

Writes the reference's NPZ schema (same keys, shapes and dtypes) from the
solvers' accessors. VTK output needs the optional ``pyevtk`` package, as in
the reference; without it a VTK request raises ImportError.
"""

from __future__ import annotations

from pathlib import Path
from typing import TYPE_CHECKING

import numpy as np

from .config import OutputConfig

if TYPE_CHECKING:
    from .simulation import FSISimulation


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

    def write_snapshot(self, sim: "FSISimulation") -> list[Path]:
        fmt = self.config.output_format
        if fmt == "npz":
            return [self.write_npz_snapshot(sim)]
        if fmt == "vtk":
            return self.write_vtk_snapshot(sim)
        if fmt == "both":
            return [self.write_npz_snapshot(sim), *self.write_vtk_snapshot(sim)]
        raise ValueError(f"Unsupported output format: {fmt}.")

    def write_npz_snapshot(self, sim: "FSISimulation") -> Path:
        # fsi/output.py:39-116
        self.ensure_output_dir()
        path = self.output_dir / f"snapshot_{sim.step_index:06d}.npz"
        arrays: dict[str, np.ndarray] = {
            "step": np.array(sim.step_index, dtype=np.int64),
            "time": np.array(sim.time, dtype=np.float64),
            "lbm_shape": np.array([sim.lbm.nx, sim.lbm.ny, sim.lbm.nz], dtype=np.int32),
            "mpm_particle_count": np.array(sim.mpm.particle_count(), dtype=np.int32),
        }
        if self.config.write_lbm_fields:
            arrays.update(
                {
                    "lbm_density": sim.lbm.density_numpy(),
                    "lbm_velocity": sim.lbm.velocity_numpy(),
                    "lbm_force": sim.lbm.force_numpy(),
                    "lbm_solid": sim.lbm.solid_numpy(),
                }
            )
        if self.config.write_mpm_particles:
            arrays.update(
                {
                    "mpm_positions": sim.mpm.positions_numpy(),
                    "mpm_velocities": sim.mpm.velocities_numpy(),
                    "mpm_particle_forces": sim.mpm.particle_forces_numpy(),
                    "mpm_deformation_gradients": sim.mpm.deformation_gradients_numpy(),
                    "mpm_active": sim.mpm.active_numpy(),
                }
            )
        if self.config.write_coupling_fields:
            cd = sim.coupler.coupling_diagnostics()
            arrays.update(
                {
                    "coupling_force": sim.coupler.coupling_force_numpy(),
                    "solid_volume_fraction": sim.coupler.solid_volume_fraction_numpy(),
                    "immersed_boundary_force": sim.coupler.immersed_boundary_force_numpy(),
                    "total_particle_coupling_force": sim.coupler.total_particle_coupling_force(),
                    "total_fluid_coupling_force": sim.coupler.total_fluid_coupling_force(),
                    "coupling_particle_valid_weight": sim.coupler.particle_valid_weights_numpy(),
                    "coupling_particle_mask": sim.coupler.particle_coupling_mask_numpy(),
                    "particle_contact_mask": sim.coupler.particle_contact_mask_numpy(),
                    "ib_total_force": np.array(
                        [cd["ib_total_force_x"], cd["ib_total_force_y"], cd["ib_total_force_z"]],
                        dtype=np.float32,
                    ),
                    "ib_active_cell_count": np.array(cd["ib_active_cell_count"], dtype=np.int32),
                    "ib_clipped_cell_count": np.array(cd["ib_clipped_cell_count"], dtype=np.int32),
                    "contact_candidate_count": np.array(cd["contact_candidate_count"], dtype=np.int32),
                    "contact_damped_particle_count": np.array(
                        cd["contact_damped_particle_count"], dtype=np.int32
                    ),
                    "coupling_unsupported_particle_count": np.array(
                        cd["unsupported_particle_count"], dtype=np.int32
                    ),
                    "coupling_partial_support_particle_count": np.array(
                        cd["partial_support_particle_count"], dtype=np.int32
                    ),
                    "coupling_clipped_particle_count": np.array(
                        cd["clipped_particle_count"], dtype=np.int32
                    ),
                }
            )
        np.savez_compressed(path, **arrays)
        return path

    def write_vtk_snapshot(self, sim: "FSISimulation") -> list[Path]:
        # fsi/output.py:118-196 (needs pyevtk, an optional dependency)
        from pyevtk.hl import imageToVTK, pointsToVTK  # noqa: F401  (ImportError when absent)

        self.ensure_output_dir()
        step = sim.step_index
        grid_base = self.output_dir / f"lbm_{step:06d}"
        particle_base = self.output_dir / f"mpm_{step:06d}"
        velocity = sim.lbm.velocity_numpy()
        cell_data = {
            "density": np.ascontiguousarray(sim.lbm.density_numpy()),
            "solid": np.ascontiguousarray(sim.lbm.solid_numpy().astype(np.float32)),
            "velocity": tuple(np.ascontiguousarray(velocity[..., a]) for a in range(3)),
        }
        grid_path = imageToVTK(str(grid_base), cellData=cell_data)
        pos = sim.mpm.positions_numpy()
        vel = sim.mpm.velocities_numpy()
        point_path = pointsToVTK(
            str(particle_base),
            np.ascontiguousarray(pos[:, 0]), np.ascontiguousarray(pos[:, 1]),
            np.ascontiguousarray(pos[:, 2]),
            data={"velocity_x": np.ascontiguousarray(vel[:, 0]),
                  "velocity_y": np.ascontiguousarray(vel[:, 1]),
                  "velocity_z": np.ascontiguousarray(vel[:, 2])},
        )
        return [Path(grid_path), Path(point_path)]
