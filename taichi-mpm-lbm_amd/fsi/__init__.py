"""Coupled LBM-MPM fluid-solid time step on AMD MI355X (gfx950).

Drop-in for the hot path of the reference package ``fsi``
(lizhuoh9/taichi-mpm-lbm): the same config dataclasses and solver classes,
with state in PyTorch-ROCm device tensors and every kernel a hand-written HIP
kernel of ``_lib/libfsi_hip.so`` called through a C ABI.

Solver classes are imported lazily (reference ``fsi/__init__.py:65-151``), so
``import fsi`` only needs the config module; touching a solver loads torch
and the kernel library.
"""

from .config import (
    BoundaryConfig,
    CouplingConfig,
    LBMConfig,
    MPMConfig,
    OutputConfig,
    SimulationConfig,
)

__all__ = [
    "BoundaryConfig",
    "CouplingConfig",
    "FSISimulation",
    "LBMConfig",
    "LBMMpmCoupler",
    "LBMSolver3D",
    "MPMConfig",
    "MPMSolver3D",
    "OutputConfig",
    "ReferenceDataset",
    "ReferenceReport",
    "SimulationConfig",
    "SimulationOutputWriter",
    "SnapshotInfo",
    "ValidationMetric",
    "ValidationReport",
    "extract_snapshot_timeseries",
    "list_npz_snapshots",
    "load_reference_dataset",
    "run_reference_validation_suite",
    "run_validation_suite",
]

# name -> (module, attribute); the solver modules load torch and the kernel
# library, the others only numpy
_LAZY = {
    "LBMSolver3D": ("lbm3d", "LBMSolver3D"),
    "MPMSolver3D": ("mpm3d", "MPMSolver3D"),
    "LBMMpmCoupler": ("coupling", "LBMMpmCoupler"),
    "FSISimulation": ("simulation", "FSISimulation"),
    "SimulationOutputWriter": ("output", "SimulationOutputWriter"),
    "SnapshotInfo": ("postprocess", "SnapshotInfo"),
    "extract_snapshot_timeseries": ("postprocess", "extract_snapshot_timeseries"),
    "list_npz_snapshots": ("postprocess", "list_npz_snapshots"),
    "ReferenceDataset": ("reference", "ReferenceDataset"),
    "ReferenceReport": ("reference", "ReferenceReport"),
    "load_reference_dataset": ("reference", "load_reference_dataset"),
    "ValidationMetric": ("validation", "ValidationMetric"),
    "ValidationReport": ("validation", "ValidationReport"),
    "run_validation_suite": ("validation_cases", "run_validation_suite"),
    "run_reference_validation_suite": ("reference_cases", "run_reference_validation_suite"),
}


def __getattr__(name: str):
    if name in _LAZY:
        import importlib

        module_name, attr = _LAZY[name]
        value = getattr(importlib.import_module(f".{module_name}", __name__), attr)
        globals()[name] = value
        return value
    raise AttributeError(f"module {__name__!r} has no attribute {name!r}")
