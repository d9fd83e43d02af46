"""Configuration dataclasses of the coupled LBM-MPM step.

Same classes, field names, defaults and validation rules as the reference's
``fsi/config.py:20-259`` so existing scene scripts construct the same objects.
Two build-only fields are added with defaults that keep the reference
behaviour: ``LBMConfig.storage`` selects the population storage format
("fp32" as in the reference, "fp16" shifted/scaled IEEE half storage, "i16"
shifted signed-normalised 16-bit storage with an adaptive scale, "fp64" the
double-precision build), and ``MPMConfig``/``CouplingConfig`` are unchanged.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path
from typing import Literal

Backend = Literal["cpu", "gpu", "cuda", "vulkan"]
BoundaryType = Literal["periodic", "pressure", "velocity", "wall"]
CouplingKernel = Literal["quadratic"]
OutputFormat = Literal["npz", "vtk", "both"]
PopulationStorage = Literal["fp32", "fp16", "i16", "fp64"]

_BACKENDS = ("cpu", "gpu", "cuda", "vulkan")
_BOUNDARY_TYPES = ("periodic", "pressure", "velocity", "wall")
_COUPLING_KERNELS = ("quadratic",)
_OUTPUT_FORMATS = ("npz", "vtk", "both")
_MATERIALS = ("elastic",)
_STORAGES = ("fp32", "fp16", "i16", "fp64")


def _require(condition: bool, message: str) -> None:
    if not condition:
        raise ValueError(message)


@dataclass(frozen=True)
class BoundaryConfig:
    """One x/y/z face of the lattice (reference ``config.py:20-34``)."""

    boundary_type: BoundaryType = "periodic"
    rho: float = 1.0
    velocity: tuple[float, float, float] = (0.0, 0.0, 0.0)

    def validate(self) -> None:
        _require(self.boundary_type in _BOUNDARY_TYPES,
                 f"Unsupported boundary type: {self.boundary_type}.")
        _require(self.rho > 0.0, f"Boundary rho must be positive, got {self.rho}.")
        _require(len(self.velocity) == 3, "Boundary velocity must have 3 components.")


@dataclass(frozen=True)
class LBMConfig:
    """D3Q19 MRT lattice settings (reference ``config.py:37-83``)."""

    nx: int = 64
    ny: int = 32
    nz: int = 16

    viscosity: float = 0.1
    rho0: float = 1.0

    force: tuple[float, float, float] = (0.0, 0.0, 0.0)
    initial_velocity: tuple[float, float, float] = (0.0, 0.0, 0.0)

    x_left: BoundaryConfig = field(default_factory=lambda: BoundaryConfig("velocity"))
    x_right: BoundaryConfig = field(default_factory=lambda: BoundaryConfig("pressure"))
    y_left: BoundaryConfig = field(default_factory=BoundaryConfig)
    y_right: BoundaryConfig = field(default_factory=BoundaryConfig)
    z_left: BoundaryConfig = field(default_factory=BoundaryConfig)
    z_right: BoundaryConfig = field(default_factory=BoundaryConfig)

    use_sparse: bool = False
    # population storage on the device: "fp32" (reference arithmetic),
    # "fp16" (f - w*rho0 scaled by a power of two, FP32 collision), "i16"
    # ((f - w*rho0)/w as 16-bit integers of an adaptive range) or "fp64"
    storage: PopulationStorage = "fp32"

    def validate(self) -> None:
        dims = (self.nx, self.ny, self.nz)
        _require(min(dims) > 4,
                 f"LBM grid is too small: {dims}. Each dimension should be greater than 4.")
        _require(self.viscosity > 0.0, f"LBM viscosity must be positive, got {self.viscosity}.")
        _require(self.rho0 > 0.0, f"LBM rho0 must be positive, got {self.rho0}.")
        _require(len(self.force) == 3, "LBM force must have 3 components.")
        _require(len(self.initial_velocity) == 3, "LBM initial_velocity must have 3 components.")
        _require(self.storage in _STORAGES, f"Unsupported population storage: {self.storage}.")
        for face in (self.x_left, self.x_right, self.y_left, self.y_right, self.z_left, self.z_right):
            face.validate()


@dataclass(frozen=True)
class MPMConfig:
    """MLS-MPM solid settings (reference ``config.py:86-138``)."""

    nx: int = 64
    ny: int = 32
    nz: int = 16

    max_particles: int = 200_000
    particles_per_cell: int = 2

    dx: float = 1.0
    dt: float = 0.25

    density: float = 1.0
    youngs_modulus: float = 1_000.0
    poisson_ratio: float = 0.25

    gravity: tuple[float, float, float] = (0.0, -9.8e-4, 0.0)

    apic: bool = True
    material: Literal["elastic"] = "elastic"
    boundary_width: int = 3
    boundary_damping: float = 0.0

    def validate(self) -> None:
        dims = (self.nx, self.ny, self.nz)
        _require(min(dims) > 4,
                 f"MPM grid is too small: {dims}. Each dimension should be greater than 4.")
        _require(self.max_particles > 0, "max_particles must be positive.")
        _require(self.particles_per_cell > 0, "particles_per_cell must be positive.")
        _require(self.dx > 0.0, "MPM dx must be positive.")
        _require(self.dt > 0.0, "MPM dt must be positive.")
        _require(self.density > 0.0, "MPM density must be positive.")
        _require(self.youngs_modulus > 0.0, "Young's modulus must be positive.")
        _require(-1.0 < self.poisson_ratio < 0.5,
                 f"Poisson ratio must be in (-1, 0.5), got {self.poisson_ratio}.")
        _require(len(self.gravity) == 3, "MPM gravity must have 3 components.")
        _require(self.material in _MATERIALS, f"Unsupported MPM material: {self.material}.")
        _require(self.boundary_width >= 1, "boundary_width must be at least 1.")
        _require(0.0 <= self.boundary_damping <= 1.0, "boundary_damping must be in [0, 1].")


@dataclass(frozen=True)
class CouplingConfig:
    """Penalty coupling settings (reference ``config.py:141-196``).

    F_p = gamma * m_p * (u_fluid_at_particle - v_particle)
    """

    enabled: bool = True
    gamma: float = 0.05
    mpm_substeps_per_lbm_step: int = 4
    kernel: CouplingKernel = "quadratic"

    build_solid_volume_fraction: bool = True
    force_limit: float | None = None
    relative_velocity_limit: float | None = None
    gamma_ramp_steps: int = 0
    min_valid_weight: float = 1.0e-6
    immersed_boundary_enabled: bool = False
    immersed_boundary_drag: float = 0.0
    immersed_boundary_fraction_threshold: float = 0.1
    immersed_boundary_max_force: float | None = None
    contact_enabled: bool = False
    contact_velocity_damping: float = 0.0
    contact_fraction_threshold: float = 0.5
    # MPM cells per lattice cell along each axis (no counterpart in the
    # reference, whose grids are one, config.py:246-255): r > 1 couples an
    # MPM grid r times finer than the lattice (MPM dims = r x LBM dims, one
    # lattice cell = r MPM dx); the coupling and contact stencils then use
    # the lattice's cells
    mpm_refinement: int = 1

    def validate(self) -> None:
        _require(self.gamma >= 0.0, f"Coupling gamma must be non-negative, got {self.gamma}.")
        _require(isinstance(self.mpm_refinement, int) and self.mpm_refinement >= 1,
                 f"mpm_refinement must be a positive integer, got {self.mpm_refinement}.")
        _require(self.mpm_substeps_per_lbm_step > 0, "mpm_substeps_per_lbm_step must be positive.")
        _require(self.kernel in _COUPLING_KERNELS, f"Unsupported coupling kernel: {self.kernel}.")
        _require(self.force_limit is None or self.force_limit > 0.0,
                 "force_limit must be positive when provided.")
        _require(self.relative_velocity_limit is None or self.relative_velocity_limit > 0.0,
                 "relative_velocity_limit must be positive when provided.")
        _require(self.gamma_ramp_steps >= 0, "gamma_ramp_steps must be non-negative.")
        _require(self.min_valid_weight > 0.0, "min_valid_weight must be positive.")
        _require(self.min_valid_weight <= 1.0, "min_valid_weight must not exceed 1.")
        _require(self.immersed_boundary_drag >= 0.0, "immersed_boundary_drag must be non-negative.")
        _require(0.0 <= self.immersed_boundary_fraction_threshold <= 1.0,
                 "immersed_boundary_fraction_threshold must be in [0, 1].")
        _require(self.immersed_boundary_max_force is None or self.immersed_boundary_max_force > 0.0,
                 "immersed_boundary_max_force must be positive when provided.")
        _require(0.0 <= self.contact_velocity_damping <= 1.0,
                 "contact_velocity_damping must be in [0, 1].")
        _require(0.0 <= self.contact_fraction_threshold <= 1.0,
                 "contact_fraction_threshold must be in [0, 1].")


@dataclass(frozen=True)
class OutputConfig:
    """Snapshot settings (reference ``config.py:199-215``)."""

    output_dir: Path = Path("outputs")
    output_interval: int = 100
    output_format: OutputFormat = "npz"

    write_lbm_fields: bool = True
    write_mpm_particles: bool = True
    write_coupling_fields: bool = True

    def validate(self) -> None:
        _require(self.output_interval > 0, "output_interval must be positive.")
        _require(self.output_format in _OUTPUT_FORMATS,
                 f"Unsupported output format: {self.output_format}.")


@dataclass(frozen=True)
class SimulationConfig:
    """Top-level settings (reference ``config.py:218-259``)."""

    backend: Backend = "cpu"
    random_seed: int = 0

    num_steps: int = 1000
    lbm_dt: float = 1.0

    lbm: LBMConfig = field(default_factory=LBMConfig)
    mpm: MPMConfig = field(default_factory=MPMConfig)
    coupling: CouplingConfig = field(default_factory=CouplingConfig)
    output: OutputConfig = field(default_factory=OutputConfig)

    def validate(self) -> None:
        _require(self.backend in _BACKENDS, f"Unsupported backend: {self.backend}.")
        _require(self.num_steps > 0, "num_steps must be positive.")
        _require(self.lbm_dt > 0.0, "lbm_dt must be positive.")
        self.lbm.validate()
        self.mpm.validate()
        self.coupling.validate()
        self.output.validate()
        lbm_dims = (self.lbm.nx, self.lbm.ny, self.lbm.nz)
        mpm_dims = (self.mpm.nx, self.mpm.ny, self.mpm.nz)
        r = self.coupling.mpm_refinement
        if r == 1:
            _require(lbm_dims == mpm_dims,
                     "For MVP, LBM and MPM grids must have the same resolution. "
                     f"Got LBM={lbm_dims}, MPM={mpm_dims}.")
        else:
            _require(mpm_dims == tuple(r * n for n in lbm_dims),
                     f"With mpm_refinement={r} the MPM grid must be {r} x the LBM grid. "
                     f"Got LBM={lbm_dims}, MPM={mpm_dims}.")

    @property
    def mpm_dt(self) -> float:
        return self.lbm_dt / float(self.coupling.mpm_substeps_per_lbm_step)
