"""Snapshot writer (reference: fsi/output.py:15-198).

Both formats are built from one set of raw fields (the solvers' accessor
arrays) through field tables: one row per output array, naming the output
group that switches it on (``OutputConfig.write_lbm_fields``,
``write_mpm_particles``, ``write_coupling_fields``), the key under which it
is written, and how it derives from the raw fields. The NPZ keys, dtypes and
shapes and the VTK file names, array names and grid spacing are the
reference's, so readers of its snapshots (fsi/postprocess.py) keep working.

``write_snapshot`` reads the raw fields with the solvers' ``*_numpy``
accessors (each a device-to-host copy). ``write_snapshot_async`` takes the
same fields from the ``*_device`` accessors, copies them on the device
(the simulation may step on at once), moves the copies to pinned host memory
on a separate stream and writes the files on a worker thread; it returns a
``concurrent.futures.Future`` of the written paths. The files are the same.

VTK needs the optional ``pyevtk`` package, as in the reference. It is
imported when a VTK snapshot is written, so the NPZ path works without it.
"""

from __future__ import annotations

from concurrent.futures import Future, ThreadPoolExecutor
from pathlib import Path
from typing import TYPE_CHECKING, Callable

import numpy as np

from .config import OutputConfig

if TYPE_CHECKING:
    from .simulation import FSISimulation

_LBM, _MPM, _COUPLING = "write_lbm_fields", "write_mpm_particles", "write_coupling_fields"

# Raw fields: (name, groups that need it, host reader, device reader).
_RAW_FIELDS: tuple[tuple[str, tuple[str, ...], Callable, Callable], ...] = (
    ("rho", (_LBM,), lambda s: s.lbm.density_numpy(), lambda s: s.lbm.density_device()),
    ("velocity", (_LBM,), lambda s: s.lbm.velocity_numpy(), lambda s: s.lbm.velocity_device()),
    ("force", (_LBM,), lambda s: s.lbm.force_numpy(), lambda s: s.lbm.force_device()),
    ("solid", (_LBM, _COUPLING), lambda s: s.lbm.solid_numpy(), lambda s: s.lbm.solid_device()),
    ("positions", (_MPM,), lambda s: s.mpm.positions_numpy(), lambda s: s.mpm.positions_device()),
    ("velocities", (_MPM,), lambda s: s.mpm.velocities_numpy(), lambda s: s.mpm.velocities_device()),
    ("particle_forces", (_MPM, _COUPLING), lambda s: s.mpm.particle_forces_numpy(),
     lambda s: s.mpm.particle_forces_device()),
    ("deformation_gradients", (_MPM,), lambda s: s.mpm.deformation_gradients_numpy(),
     lambda s: s.mpm.deformation_gradients_device()),
    ("active", (_MPM,), lambda s: s.mpm.active_numpy(), lambda s: s.mpm.active_device()),
    ("coupling_force", (_COUPLING,), lambda s: s.coupler.coupling_force_numpy(),
     lambda s: s.coupler.coupling_force_device()),
    ("solid_volume_fraction", (_COUPLING,), lambda s: s.coupler.solid_volume_fraction_numpy(),
     lambda s: s.coupler.solid_volume_fraction_device()),
    ("immersed_boundary_force", (_COUPLING,), lambda s: s.coupler.immersed_boundary_force_numpy(),
     lambda s: s.coupler.immersed_boundary_force_device()),
    ("valid_weight", (_COUPLING,), lambda s: s.coupler.particle_valid_weights_numpy(),
     lambda s: s.coupler.particle_valid_weights_device()),
    ("coupling_mask", (_COUPLING,), lambda s: s.coupler.particle_coupling_mask_numpy(),
     lambda s: s.coupler.particle_coupling_mask_device()),
    ("contact_mask", (_COUPLING,), lambda s: s.coupler.particle_contact_mask_numpy(),
     lambda s: s.coupler.particle_contact_mask_device()),
)


def _counter(key: str) -> Callable:
    """An integer of coupling_diagnostics() as a 0-d int32 array."""
    return lambda raw, sc: np.array(sc["diag"][key], dtype=np.int32)


# NPZ snapshot, fsi/output.py:39-116: (group or None for always, key,
# builder). Builders take the raw host fields and the scalars captured with
# them (step, time, shape, particle count, coupling_diagnostics()). The two
# totals are the reference's float32 host sums (fsi/coupling.py:238-244).
_NPZ_FIELDS: tuple[tuple[str | None, str, Callable], ...] = (
    (None, "step", lambda raw, sc: np.array(sc["step"], dtype=np.int64)),
    (None, "time", lambda raw, sc: np.array(sc["time"], dtype=np.float64)),
    (None, "lbm_shape", lambda raw, sc: np.array(sc["shape"], dtype=np.int32)),
    (None, "mpm_particle_count", lambda raw, sc: np.array(sc["count"], dtype=np.int32)),
    (_LBM, "lbm_density", lambda raw, sc: raw["rho"]),
    (_LBM, "lbm_velocity", lambda raw, sc: raw["velocity"]),
    (_LBM, "lbm_force", lambda raw, sc: raw["force"]),
    (_LBM, "lbm_solid", lambda raw, sc: raw["solid"]),
    (_MPM, "mpm_positions", lambda raw, sc: raw["positions"]),
    (_MPM, "mpm_velocities", lambda raw, sc: raw["velocities"]),
    (_MPM, "mpm_particle_forces", lambda raw, sc: raw["particle_forces"]),
    (_MPM, "mpm_deformation_gradients", lambda raw, sc: raw["deformation_gradients"]),
    (_MPM, "mpm_active", lambda raw, sc: raw["active"]),
    (_COUPLING, "coupling_force", lambda raw, sc: raw["coupling_force"]),
    (_COUPLING, "solid_volume_fraction", lambda raw, sc: raw["solid_volume_fraction"]),
    (_COUPLING, "immersed_boundary_force", lambda raw, sc: raw["immersed_boundary_force"]),
    (_COUPLING, "total_particle_coupling_force", lambda raw, sc: raw["particle_forces"].sum(axis=0)),
    (_COUPLING, "total_fluid_coupling_force",
     lambda raw, sc: raw["coupling_force"][raw["solid"] == 0].sum(axis=0)),
    (_COUPLING, "coupling_particle_valid_weight", lambda raw, sc: raw["valid_weight"]),
    (_COUPLING, "coupling_particle_mask", lambda raw, sc: raw["coupling_mask"]),
    (_COUPLING, "particle_contact_mask", lambda raw, sc: raw["contact_mask"]),
    (_COUPLING, "ib_total_force",
     lambda raw, sc: np.array([sc["diag"][f"ib_total_force_{a}"] for a in "xyz"], dtype=np.float32)),
    (_COUPLING, "ib_active_cell_count", _counter("ib_active_cell_count")),
    (_COUPLING, "ib_clipped_cell_count", _counter("ib_clipped_cell_count")),
    (_COUPLING, "contact_candidate_count", _counter("contact_candidate_count")),
    (_COUPLING, "contact_damped_particle_count", _counter("contact_damped_particle_count")),
    (_COUPLING, "coupling_unsupported_particle_count", _counter("unsupported_particle_count")),
    (_COUPLING, "coupling_partial_support_particle_count", _counter("partial_support_particle_count")),
    (_COUPLING, "coupling_clipped_particle_count", _counter("clipped_particle_count")),
)

# VTK cell data of fluid_%06d.vti, fsi/output.py:128-166: (group, name, raw
# field, vector). Vectors are written as their three component arrays.
_VTK_CELL_FIELDS: tuple[tuple[str, str, str, bool], ...] = (
    (_LBM, "rho", "rho", False),
    (_LBM, "velocity", "velocity", True),
    (_LBM, "force", "force", True),
    (_LBM, "solid", "solid", False),
    (_COUPLING, "coupling_force", "coupling_force", True),
    (_COUPLING, "solid_volume_fraction", "solid_volume_fraction", False),
    (_COUPLING, "immersed_boundary_force", "immersed_boundary_force", True),
)

# VTK point data of particles_%06d.vtu, fsi/output.py:168-189
_VTK_POINT_FIELDS: tuple[tuple[str | None, str, str, bool], ...] = (
    (None, "velocity", "velocities", True),
    (None, "particle_force", "particle_forces", True),
    (_COUPLING, "contact_mask", "contact_mask", False),
)


def _vtk_array(values: np.ndarray, vector: bool):
    if vector:
        return tuple(np.ascontiguousarray(values[..., a]) for a in range(3))
    return np.ascontiguousarray(values)


class SimulationOutputWriter:
    """Snapshot writer (reference: fsi/output.py): NPZ files of the lattice,
    coupling and particle fields with the reference's keys, and VTK files
    when pyevtk is importable, synchronously or from a worker thread."""

    def __init__(self, config: OutputConfig):
        self.config = config
        self.config.validate()
        self.output_dir = Path(config.output_dir)
        self._pool: ThreadPoolExecutor | None = None

    def ensure_output_dir(self) -> None:
        self.output_dir.mkdir(parents=True, exist_ok=True)

    def should_write(self, step: int) -> bool:
        return int(step) % self.config.output_interval == 0

    def _enabled(self, group: str | None) -> bool:
        return group is None or bool(getattr(self.config, group))

    def _check_format(self) -> None:
        if self.config.output_format not in ("npz", "vtk", "both"):
            raise ValueError(f"Unsupported output format: {self.config.output_format}.")

    # ------------------------------------------------------ collection --

    def _wanted(self) -> list[tuple[str, Callable, Callable]]:
        fmt = self.config.output_format
        groups = {g for g in (_LBM, _MPM, _COUPLING) if self._enabled(g)}
        if fmt in ("vtk", "both") and self.config.write_mpm_particles:
            groups.add(_MPM)  # the point cloud's positions and point data
        return [(name, host, dev) for name, need, host, dev in _RAW_FIELDS if groups.intersection(need)]

    def _scalars(self, sim: "FSISimulation") -> dict:
        return {
            "step": int(sim.step_index), "time": float(sim.time),
            "shape": [sim.lbm.nx, sim.lbm.ny, sim.lbm.nz], "count": int(sim.mpm.particle_count()),
            "dx": float(sim.mpm.dx),
            "diag": sim.coupler.coupling_diagnostics() if self.config.write_coupling_fields else None,
        }

    # --------------------------------------------------------- writing --

    def write_snapshot(self, sim: "FSISimulation") -> list[Path]:
        self._check_format()
        raw = {name: host(sim) for name, host, _ in self._wanted()}
        return self._emit(raw, self._scalars(sim))

    def write_snapshot_async(self, sim: "FSISimulation") -> Future:
        """Start a snapshot of the current state and return at once; the
        Future gives the written paths. The device copies are taken on the
        current stream before this returns, so the simulation may keep
        stepping; the host copies run on a separate stream into pinned
        memory and the files are written by a worker thread."""
        import torch

        self._check_format()
        scalars = self._scalars(sim)
        compute = torch.cuda.current_stream()
        frozen = {name: dev(sim).clone() for name, _, dev in self._wanted()}
        copy = torch.cuda.Stream()
        copy.wait_stream(compute)
        host = {}
        with torch.cuda.stream(copy):
            for name, t in frozen.items():
                h = torch.empty(t.shape, dtype=t.dtype, device="cpu", pin_memory=True)
                h.copy_(t, non_blocking=True)
                t.record_stream(copy)
                host[name] = h
            done = torch.cuda.Event()
            done.record(copy)
        del frozen
        if self._pool is None:
            self._pool = ThreadPoolExecutor(max_workers=1, thread_name_prefix="fsi-snapshot")

        def finish() -> list[Path]:
            done.synchronize()
            return self._emit({name: h.numpy() for name, h in host.items()}, scalars)

        return self._pool.submit(finish)

    def _emit(self, raw: dict, scalars: dict) -> list[Path]:
        writers = {
            "npz": lambda: [self._write_npz(raw, scalars)],
            "vtk": lambda: self._write_vtk(raw, scalars),
            "both": lambda: [self._write_npz(raw, scalars), *self._write_vtk(raw, scalars)],
        }
        return writers[self.config.output_format]()

    def write_npz_snapshot(self, sim: "FSISimulation") -> Path:
        raw = {name: host(sim) for name, host, _ in self._wanted()}
        return self._write_npz(raw, self._scalars(sim))

    def write_vtk_snapshot(self, sim: "FSISimulation") -> list[Path]:
        raw = {name: host(sim) for name, host, _ in self._wanted()}
        return self._write_vtk(raw, self._scalars(sim))

    def _write_npz(self, raw: dict, sc: dict) -> Path:
        self.ensure_output_dir()
        path = self.output_dir / f"snapshot_{sc['step']:06d}.npz"
        arrays = {key: build(raw, sc) for group, key, build in _NPZ_FIELDS if self._enabled(group)}
        np.savez_compressed(path, **arrays)
        return path

    def _write_vtk(self, raw: dict, sc: dict) -> list[Path]:
        from pyevtk.hl import imageToVTK, pointsToVTK  # ImportError when pyevtk is absent

        self.ensure_output_dir()
        step = sc["step"]
        paths: list[Path] = []
        if self.config.write_lbm_fields or self.config.write_coupling_fields:
            cells = {name: _vtk_array(raw[field], vec) for group, name, field, vec in _VTK_CELL_FIELDS
                     if self._enabled(group)}
            dx = sc["dx"]
            paths.append(Path(imageToVTK(str(self.output_dir / f"fluid_{step:06d}"), origin=(0.0, 0.0, 0.0),
                                         spacing=(dx, dx, dx), cellData=cells)))
        if self.config.write_mpm_particles:
            pos = raw["positions"]
            xyz = [np.ascontiguousarray(pos[:, a]) if len(pos) else np.zeros(0, np.float32) for a in range(3)]
            points = {name: _vtk_array(raw[field], vec) for group, name, field, vec in _VTK_POINT_FIELDS
                      if self._enabled(group)}
            paths.append(Path(pointsToVTK(str(self.output_dir / f"particles_{step:06d}"), *xyz, data=points)))
        return paths
