"""CPU oracle objects mirroring the reference's solver API.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` as the checker. The product
package never imports it.

Each class restates one reference class (state held in numpy arrays with the
reference's Taichi layouts, kernels in ``fsi_oracle.cpp``):

* ``OracleLBM``       -> fsi/lbm3d.py:57-534  (LBMSolver3D)
* ``OracleMPM``       -> fsi/mpm3d.py:9-368   (MPMSolver3D)
* ``OracleCoupler``   -> fsi/coupling.py:11-496 (LBMMpmCoupler)
* ``OracleSimulation``-> fsi/simulation.py:12-153 (FSISimulation)

``precision`` selects the float32 instantiation (the reference's arithmetic
type) or the float64 one (the high-precision oracle for error statements).
Config objects are duck-typed: anything with the reference config attributes.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "liboracle.so"
_lib = None

Q = 19
E = np.array(
    [[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
     [1, 1, 0], [-1, -1, 0], [1, -1, 0], [-1, 1, 0], [1, 0, 1], [-1, 0, -1], [1, 0, -1],
     [-1, 0, 1], [0, 1, 1], [0, -1, -1], [0, 1, -1], [0, -1, 1]],
    dtype=np.int32,
)
LR = np.array([0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17], dtype=np.int32)

# fsi/lbm3d.py:30-54: the MRT transform, 19 rows of small integers.
MRT_MATRIX = np.array(
    [
        [1] * 19,
        [-1, 0, 0, 0, 0, 0, 0] + [1] * 12,
        [1, -2, -2, -2, -2, -2, -2] + [1] * 12,
        [0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0],
        [0, -2, 2, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0],
        [0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1],
        [0, 0, 0, -2, 2, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1],
        [0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1],
        [0, 0, 0, 0, 0, -2, 2, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1],
        [0, 2, 2, -1, -1, -1, -1, 1, 1, 1, 1, 1, 1, 1, 1, -2, -2, -2, -2],
        [0, -2, -2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, -2, -2, -2, -2],
        [0, 0, 0, 1, 1, -1, -1, 1, 1, 1, 1, -1, -1, -1, -1, 0, 0, 0, 0],
        [0, 0, 0, -1, -1, 1, 1, 1, 1, 1, 1, -1, -1, -1, -1, 0, 0, 0, 0],
        [0, 0, 0, 0, 0, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0],
        [0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, -1, -1],
        [0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0],
        [0, 0, 0, 0, 0, 0, 0, 1, -1, 1, -1, -1, 1, -1, 1, 0, 0, 0, 0],
        [0, 0, 0, 0, 0, 0, 0, -1, 1, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1],
        [0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, -1, -1, 1, -1, 1, 1, -1],
    ],
    dtype=np.float64,
)


def build(force: bool = False) -> Path:
    """Compile fsi_oracle.cpp into liboracle.so (recipe also in oracle/Makefile)."""
    src = _HERE / "fsi_oracle.cpp"
    if force or not _LIB_PATH.exists() or _LIB_PATH.stat().st_mtime < src.stat().st_mtime:
        subprocess.run(["make", "-C", str(_HERE), "-s", "liboracle.so"], check=True)
    return _LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        _lib = ctypes.CDLL(str(_LIB_PATH))
    return _lib


class MpmArgs(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("num_particles", ctypes.c_int),
        ("dx", ctypes.c_double), ("inv_dx", ctypes.c_double), ("dt", ctypes.c_double),
        ("two_mu", ctypes.c_double), ("lam", ctypes.c_double),
        ("gravity", ctypes.c_double * 3),
        ("boundary_width", ctypes.c_int), ("boundary_damping", ctypes.c_double),
        ("apic", ctypes.c_int), ("four_inv_dx", ctypes.c_double),
        ("clamp_lower", ctypes.c_double), ("clamp_upper", ctypes.c_double * 3),
    ]


class CouplingArgs(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("num_particles", ctypes.c_int),
        ("inv_dx", ctypes.c_double), ("gamma", ctypes.c_double), ("dt_ratio", ctypes.c_double),
        ("cell_volume", ctypes.c_double), ("build_fraction", ctypes.c_int),
        ("min_valid_weight", ctypes.c_double),
        ("use_relative_velocity_limit", ctypes.c_int), ("relative_velocity_limit", ctypes.c_double),
        ("use_force_limit", ctypes.c_int), ("force_limit", ctypes.c_double),
        ("gamma_ramp_steps", ctypes.c_int), ("coupled_step", ctypes.c_int),
    ]


def _p(a: np.ndarray) -> ctypes.c_void_p:
    assert a.flags["C_CONTIGUOUS"]
    return ctypes.c_void_p(a.ctypes.data)


def _d3(v) -> ctypes.Array:
    return (ctypes.c_double * 3)(*[float(x) for x in v])


def _boundary_mode(b) -> int:
    # fsi/lbm3d.py:16-27
    if b.boundary_type == "periodic":
        return 0
    if b.boundary_type == "pressure":
        return 1
    if b.boundary_type == "velocity":
        return 2
    if b.boundary_type == "wall":
        raise NotImplementedError("wall face boundaries are not implemented")
    raise ValueError(f"Unsupported boundary type: {b.boundary_type}.")


def relaxation_rates(viscosity: float) -> np.ndarray:
    """fsi/lbm3d.py:141-168: S from tau = 3 nu + 0.5, evaluated in float64."""
    tau = 3.0 * viscosity + 0.5
    s_v = 1.0 / tau
    s_o = 8.0 * (2.0 - s_v) / (8.0 - s_v)
    return np.array([0.0, s_v, s_v, 0.0, s_o, 0.0, s_o, 0.0, s_o, s_v, s_v, s_v, s_v, s_v,
                     s_v, s_v, s_o, s_o, s_o])


class OracleLBM:
    """fsi/lbm3d.py:57-534 on the CPU."""

    def __init__(self, config, precision: str = "f32"):
        config.validate()
        if config.use_sparse:
            raise NotImplementedError("Sparse LBM storage is not implemented.")
        for b in (config.y_left, config.y_right, config.z_left, config.z_right):
            if b.boundary_type != "periodic":
                raise NotImplementedError("non-periodic y/z faces are not implemented")
        self.config = config
        self.precision = precision
        self.dtype = np.float32 if precision == "f32" else np.float64
        self.sfx = precision
        self.nx, self.ny, self.nz = int(config.nx), int(config.ny), int(config.nz)
        self.left_mode = _boundary_mode(config.x_left)
        self.right_mode = _boundary_mode(config.x_right)
        shape = (self.nx, self.ny, self.nz)
        dt = self.dtype
        self.f = np.zeros((*shape, Q), dt)
        self.F = np.zeros((*shape, Q), dt)
        self.moment = np.zeros((*shape, Q), dt)
        self.rho = np.zeros(shape, dt)
        self.v = np.zeros((*shape, 3), dt)
        self.force = np.zeros((*shape, 3), dt)
        self.solid = np.zeros(shape, np.int32)
        # lbm3d.py:140-174: tables in the kernel float type; inv_M = inv_f64(M) cast
        self.S = relaxation_rates(config.viscosity).astype(dt)
        self.M = MRT_MATRIX.astype(dt)
        self.invM = np.linalg.inv(MRT_MATRIX).astype(dt)
        self._initialized = False

    def _call(self, name, *args):
        return getattr(lib(), f"oref_{name}_{self.sfx}")(*args)

    def initialize(self, solid_np=None):
        if solid_np is None:
            solid = np.zeros((self.nx, self.ny, self.nz), np.int32)
        else:
            arr = np.asarray(solid_np)
            if arr.shape != (self.nx, self.ny, self.nz):
                raise ValueError("solid_np shape mismatch")
            solid = np.ascontiguousarray(np.where(arr > 0, 1, 0).astype(np.int32))
        self.solid[...] = solid
        self._call("lbm_init_fields", self.nx, self.ny, self.nz, ctypes.c_double(self.config.rho0),
                   _d3(self.config.initial_velocity), _p(self.solid), _p(self.rho), _p(self.v),
                   _p(self.f), _p(self.F))
        self.reset_force()
        self._initialized = True

    def step(self):
        if not self._initialized:
            raise RuntimeError("initialize must be called before step()")
        n = (self.nx, self.ny, self.nz)
        self._call("lbm_collide", *n, _p(self.solid), _p(self.F), _p(self.rho), _p(self.v),
                   _p(self.force), _p(self.S), _p(self.M), _p(self.invM), _p(self.moment), _p(self.f))
        self._call("lbm_stream", *n, _p(self.solid), _p(self.f), _p(self.F))
        c = self.config
        self._call("lbm_apply_boundary", *n, self.left_mode, self.right_mode,
                   ctypes.c_double(c.x_left.rho), ctypes.c_double(c.x_right.rho),
                   _d3(c.x_left.velocity), _d3(c.x_right.velocity), ctypes.c_double(c.rho0),
                   _p(self.solid), _p(self.v), _p(self.F))
        self._call("lbm_update_macro", *n, ctypes.c_double(c.rho0), _p(self.solid), _p(self.F),
                   _p(self.force), _p(self.f), _p(self.rho), _p(self.v))

    # --- force API (lbm3d.py:459-534)
    def _force_op(self, op, value=(0.0, 0.0, 0.0)):
        self._call("lbm_force_op", self.nx, self.ny, self.nz, op, _d3(value), _p(self.solid),
                   _p(self.force))

    def clear_force(self):
        self._force_op(0)

    def reset_force(self):
        self._force_op(1, self.config.force)

    def set_uniform_force(self, force):
        arr = np.asarray(force, dtype=np.float32)
        if arr.shape != (3,):
            raise ValueError("force must have 3 components.")
        self._force_op(1, arr.astype(np.float64))

    def add_uniform_force(self, force):
        arr = np.asarray(force, dtype=np.float32)
        if arr.shape != (3,):
            raise ValueError("force must have 3 components.")
        self._force_op(2, arr.astype(np.float64))

    def set_force_from_numpy(self, force_np):
        arr = np.asarray(force_np, dtype=np.float32)
        if arr.shape != (self.nx, self.ny, self.nz, 3):
            raise ValueError("force_np shape mismatch")
        self.force[...] = arr
        self.zero_force_on_solid()

    def zero_force_on_solid(self):
        self._force_op(3)

    # --- accessors (lbm3d.py:435-457, 480-483)
    def density_numpy(self):
        return self.rho.copy()

    def velocity_numpy(self):
        return self.v.copy()

    def solid_numpy(self):
        return self.solid.copy()

    def distribution_numpy(self):
        return self.f.copy()

    def force_numpy(self):
        return self.force.copy()

    def total_mass(self):
        return float(self.rho[self.solid == 0].sum())

    def max_velocity_norm(self):
        return float(np.linalg.norm(self.v, axis=-1).max())

    def total_force(self):
        return self.force[self.solid == 0].sum(axis=0)


class OracleMPM:
    """fsi/mpm3d.py:9-368 on the CPU."""

    def __init__(self, config, precision: str = "f32"):
        config.validate()
        self.config = config
        self.precision = precision
        self.dtype = np.float32 if precision == "f32" else np.float64
        self.sfx = precision
        self.nx, self.ny, self.nz = int(config.nx), int(config.ny), int(config.nz)
        self.max_particles = int(config.max_particles)
        self.dx = float(config.dx)
        self.inv_dx = 1.0 / self.dx
        self.dt_default = float(config.dt)
        self.mu = config.youngs_modulus / (2.0 * (1.0 + config.poisson_ratio))
        self.lam = config.youngs_modulus * config.poisson_ratio / (
            (1.0 + config.poisson_ratio) * (1.0 - 2.0 * config.poisson_ratio))
        dt = self.dtype
        mp = self.max_particles
        self.x = np.zeros((mp, 3), dt)
        self.v = np.zeros((mp, 3), dt)
        self.C = np.zeros((mp, 3, 3), dt)
        self.F = np.zeros((mp, 3, 3), dt)
        self.mass = np.zeros(mp, dt)
        self.volume = np.zeros(mp, dt)
        self.active = np.zeros(mp, np.int32)
        self.particle_force = np.zeros((mp, 3), dt)
        self.grid_v = np.zeros((self.nx, self.ny, self.nz, 3), dt)
        self.grid_m = np.zeros((self.nx, self.ny, self.nz), dt)
        self.num_particles = 0
        self._initialized = False

    def _call(self, name, *args):
        return getattr(lib(), f"oref_{name}_{self.sfx}")(*args)

    def initialize_particles_box(self, lower, upper, spacing=None, initial_velocity=(0.0, 0.0, 0.0)):
        lo = np.asarray(lower, np.float32)
        hi = np.asarray(upper, np.float32)
        if np.any(hi <= lo):
            raise ValueError("upper must be greater than lower in every dimension.")
        if spacing is None:
            sp = np.full(3, self.dx / float(self.config.particles_per_cell), np.float32)
        elif np.isscalar(spacing):
            sp = np.full(3, float(spacing), np.float32)
        else:
            sp = np.asarray(spacing, np.float32)
        axes = [np.arange(lo[a], hi[a], sp[a]) for a in range(3)]
        if any(len(ax) == 0 for ax in axes):
            raise ValueError("particle box produced no particles.")
        mesh = np.meshgrid(*axes, indexing="ij")
        pos = np.stack([m.ravel() for m in mesh], axis=1).astype(np.float32)
        vel = np.broadcast_to(np.asarray(initial_velocity, np.float32), pos.shape).astype(np.float32)
        return self.initialize_particles_from_numpy(pos, vel, particle_volume=float(np.prod(sp)))

    def initialize_particles_from_numpy(self, positions, velocities=None, particle_mass=None,
                                        particle_volume=None):
        pos = np.asarray(positions, np.float32)
        n = pos.shape[0]
        if n > self.max_particles:
            raise ValueError("particle count exceeds max_particles")
        vel = np.zeros((n, 3), np.float32) if velocities is None else np.asarray(velocities, np.float32)
        vol = float(particle_volume) if particle_volume is not None else \
            self.dx ** 3 / float(self.config.particles_per_cell ** 3)
        mass = float(particle_mass) if particle_mass is not None else float(self.config.density) * vol
        for arr in (self.x, self.v, self.C, self.F, self.mass, self.volume, self.active,
                    self.particle_force):
            arr[...] = 0
        self.x[:n] = pos
        self.v[:n] = vel
        self.F[:n] = np.eye(3)
        # fields are float32 in the reference: host values pass through float32
        self.mass[:n] = np.float32(mass)
        self.volume[:n] = np.float32(vol)
        self.active[:n] = 1
        self.num_particles = n
        self._initialized = True
        return n

    def args(self, dt) -> MpmArgs:
        c = self.config
        a = MpmArgs()
        a.nx, a.ny, a.nz = self.nx, self.ny, self.nz
        a.num_particles = self.num_particles
        a.dx = self.dx
        a.inv_dx = self.inv_dx
        a.dt = float(np.float32(dt)) if self.precision == "f32" else float(dt)
        a.two_mu = 2.0 * self.mu
        a.lam = self.lam
        a.gravity = (ctypes.c_double * 3)(*[float(g) for g in c.gravity])
        a.boundary_width = int(c.boundary_width)
        a.boundary_damping = float(c.boundary_damping)
        a.apic = 1 if c.apic else 0
        a.four_inv_dx = 4.0 * self.inv_dx
        a.clamp_lower = 0.5 * self.dx
        a.clamp_upper = (ctypes.c_double * 3)((self.nx - 1.5) * self.dx, (self.ny - 1.5) * self.dx,
                                              (self.nz - 1.5) * self.dx)
        return a

    def substep(self, dt=None):
        if not self._initialized:
            raise RuntimeError("particles must be initialized before substep()")
        dtv = self.dt_default if dt is None else float(dt)
        if dtv <= 0.0:
            raise ValueError("dt must be positive.")
        a = self.args(dtv)
        self.grid_v[...] = 0
        self.grid_m[...] = 0
        ap = ctypes.byref(a)
        self._call("mpm_p2g", ap, _p(self.x), _p(self.v), _p(self.C), _p(self.F), _p(self.mass),
                   _p(self.volume), _p(self.active), _p(self.particle_force), _p(self.grid_v),
                   _p(self.grid_m))
        self._call("mpm_grid_op", ap, _p(self.grid_v), _p(self.grid_m))
        self._call("mpm_g2p", ap, _p(self.x), _p(self.v), _p(self.C), _p(self.active),
                   _p(self.grid_v))

    def clear_particle_force(self):
        self.particle_force[: self.num_particles] = 0

    def particle_count(self):
        return int(self.num_particles)

    def positions_numpy(self):
        return self.x[: self.num_particles].copy()

    def velocities_numpy(self):
        return self.v[: self.num_particles].copy()

    def particle_forces_numpy(self):
        return self.particle_force[: self.num_particles].copy()

    def deformation_gradients_numpy(self):
        return self.F[: self.num_particles].copy()

    def active_numpy(self):
        return self.active[: self.num_particles].copy()

    def center_of_mass(self):
        n = self.num_particles
        if n == 0:
            return np.zeros(3, self.dtype)
        m = self.mass[:n]
        tot = m.sum()
        if tot <= 0:
            return np.zeros(3, self.dtype)
        return (self.x[:n] * m[:, None]).sum(axis=0) / tot

    def max_velocity_norm(self):
        n = self.num_particles
        if n == 0:
            return 0.0
        return float(np.linalg.norm(self.v[:n], axis=1).max())


class OracleCoupler:
    """fsi/coupling.py:11-496 on the CPU."""

    def __init__(self, config, lbm: OracleLBM, mpm: OracleMPM):
        config.validate()
        # MPM cells per lattice cell (1: the reference's co-located grids);
        # the coupling and contact kernels take the lattice's dims and cell
        # size, so a refined MPM grid needs nothing else
        r = int(getattr(config, "mpm_refinement", 1))
        if (r * lbm.nx, r * lbm.ny, r * lbm.nz) != (mpm.nx, mpm.ny, mpm.nz):
            raise ValueError("LBM and MPM grids must match.")
        self.config = config
        self.lbm, self.mpm = lbm, mpm
        self.sfx = lbm.sfx
        self.dtype = lbm.dtype
        self.enabled = bool(config.enabled)
        self.substeps = int(config.mpm_substeps_per_lbm_step)
        self.nx, self.ny, self.nz = lbm.nx, lbm.ny, lbm.nz
        self.inv_dx = 1.0 / (r * float(mpm.dx))
        self.cell_volume = (r * float(mpm.dx)) ** 3
        shape = (self.nx, self.ny, self.nz)
        dt = self.dtype
        self.coupling_force = np.zeros((*shape, 3), dt)
        self.solid_volume_fraction = np.zeros(shape, dt)
        self.immersed_boundary_force = np.zeros((*shape, 3), dt)
        mp = mpm.max_particles
        self.particle_valid_weight = np.zeros(mp, dt)
        self.particle_coupling_mask = np.zeros(mp, np.int32)
        self.particle_contact_mask = np.zeros(mp, np.int32)
        self.counters = np.zeros(8, np.int32)
        self.ib_total_force = np.zeros(3, dt)
        self.coupled_step = 0

    def _call(self, name, *args):
        return getattr(lib(), f"oref_{name}_{self.sfx}")(*args)

    def _validate_initialized(self):
        if not getattr(self.lbm, "_initialized", False):
            raise RuntimeError("LBM must be initialized before coupling.")
        if not getattr(self.mpm, "_initialized", False):
            raise RuntimeError("MPM must be initialized before coupling.")

    def clear_coupling_fields(self):
        # coupling.py:281-289 then 291-301
        self.counters[5] = self.counters[6] = 0
        self.ib_total_force[...] = 0
        self.coupling_force[...] = 0
        self.solid_volume_fraction[...] = 0
        self.immersed_boundary_force[...] = 0
        self.clear_particle_diagnostics()

    def clear_particle_diagnostics(self):
        self.counters[0:5] = 0
        self.particle_valid_weight[...] = 0
        self.particle_coupling_mask[...] = 0
        self.particle_contact_mask[...] = 0

    def _cargs(self, dt_ratio, build) -> CouplingArgs:
        c = self.config
        a = CouplingArgs()
        a.nx, a.ny, a.nz = self.nx, self.ny, self.nz
        a.num_particles = self.mpm.num_particles
        a.inv_dx = self.inv_dx
        a.gamma = float(c.gamma)
        a.dt_ratio = float(dt_ratio)
        a.cell_volume = self.cell_volume
        a.build_fraction = 1 if build else 0
        a.min_valid_weight = float(c.min_valid_weight)
        a.use_relative_velocity_limit = 0 if c.relative_velocity_limit is None else 1
        a.relative_velocity_limit = 0.0 if c.relative_velocity_limit is None else float(c.relative_velocity_limit)
        a.use_force_limit = 0 if c.force_limit is None else 1
        a.force_limit = 0.0 if c.force_limit is None else float(c.force_limit)
        a.gamma_ramp_steps = int(c.gamma_ramp_steps)
        a.coupled_step = int(self.coupled_step)
        return a

    def _coupling_kernel(self, dt_ratio, build):
        a = self._cargs(dt_ratio, build)
        m, l = self.mpm, self.lbm
        self._call("coupling_forces", ctypes.byref(a), _p(m.x), _p(m.v), _p(m.mass), _p(m.volume),
                   _p(m.active), _p(l.solid), _p(l.v), _p(m.particle_force),
                   _p(self.particle_valid_weight), _p(self.particle_coupling_mask),
                   _p(self.coupling_force), _p(self.solid_volume_fraction), _p(self.counters))

    def _contact_kernel(self):
        c = self.config
        a = self._cargs(1.0, False)
        self._call("contact", ctypes.byref(a), 1 if c.contact_enabled else 0,
                   ctypes.c_double(c.contact_fraction_threshold),
                   ctypes.c_double(c.contact_velocity_damping), _p(self.mpm.x), _p(self.mpm.v),
                   _p(self.mpm.active), _p(self.lbm.solid), _p(self.solid_volume_fraction),
                   _p(self.particle_contact_mask), _p(self.counters))

    def _ib_compute(self):
        c = self.config
        l = self.lbm
        self._call("ib_compute", self.nx, self.ny, self.nz, 1 if c.immersed_boundary_enabled else 0,
                   ctypes.c_double(c.immersed_boundary_drag),
                   ctypes.c_double(c.immersed_boundary_fraction_threshold),
                   0 if c.immersed_boundary_max_force is None else 1,
                   ctypes.c_double(0.0 if c.immersed_boundary_max_force is None else c.immersed_boundary_max_force),
                   _p(l.solid), _p(self.solid_volume_fraction), _p(l.rho), _p(l.v),
                   _p(self.immersed_boundary_force), _p(self.counters), _p(self.ib_total_force))

    def _clamp_svf(self):
        np.minimum(self.solid_volume_fraction, 1.0, out=self.solid_volume_fraction)

    def step(self, lbm_dt):
        # coupling.py:135-173
        lbm_dt = float(lbm_dt)
        if lbm_dt <= 0.0:
            raise ValueError("lbm_dt must be positive.")
        self._validate_initialized()
        self.clear_coupling_fields()
        mpm_dt = lbm_dt / float(self.substeps)
        dt_ratio = mpm_dt / lbm_dt
        for s in range(self.substeps):
            self.mpm.clear_particle_force()
            if self.enabled:
                self.clear_particle_diagnostics()
                self._coupling_kernel(dt_ratio, s == 0 and self.config.build_solid_volume_fraction)
            self.mpm.substep(dt=mpm_dt)
            if self.config.contact_enabled:
                self._contact_kernel()
        if self.enabled and self.config.build_solid_volume_fraction:
            self._clamp_svf()
        self.lbm.reset_force()
        if self.enabled:
            fl = self.lbm.solid == 0
            self.lbm.force[fl] += self.coupling_force[fl]
            self.lbm.force[~fl] = 0
        if self.config.immersed_boundary_enabled:
            self._ib_compute()
            fl = self.lbm.solid == 0
            self.lbm.force[fl] += self.immersed_boundary_force[fl]
            self.immersed_boundary_force[~fl] = 0
            self.lbm.force[~fl] = 0
        self.lbm.step()
        if self.enabled:
            self.coupled_step += 1

    def compute_coupling_forces(self, dt_ratio=1.0):
        # coupling.py:175-196
        if float(dt_ratio) < 0.0:
            raise ValueError("dt_ratio must be non-negative.")
        self._validate_initialized()
        self.clear_coupling_fields()
        self.mpm.clear_particle_force()
        if not self.enabled:
            return
        self._coupling_kernel(dt_ratio, self.config.build_solid_volume_fraction)
        if self.config.build_solid_volume_fraction:
            self._clamp_svf()
        if self.config.contact_enabled:
            self._contact_kernel()
        if self.config.immersed_boundary_enabled:
            self._ib_compute()

    # accessors (coupling.py:205-271)
    def coupling_force_numpy(self):
        return self.coupling_force.copy()

    def solid_volume_fraction_numpy(self):
        return self.solid_volume_fraction.copy()

    def immersed_boundary_force_numpy(self):
        return self.immersed_boundary_force.copy()

    def particle_valid_weights_numpy(self):
        return self.particle_valid_weight[: self.mpm.num_particles].copy()

    def particle_coupling_mask_numpy(self):
        return self.particle_coupling_mask[: self.mpm.num_particles].copy()

    def particle_contact_mask_numpy(self):
        return self.particle_contact_mask[: self.mpm.num_particles].copy()

    def total_particle_coupling_force(self):
        return self.mpm.particle_forces_numpy().sum(axis=0)

    def total_fluid_coupling_force(self):
        return self.coupling_force[self.lbm.solid == 0].sum(axis=0)

    def effective_gamma(self):
        g = float(np.float32(self.config.gamma))
        r = int(self.config.gamma_ramp_steps)
        if r <= 0:
            return g
        return g * min(1.0, float(self.coupled_step + 1) / float(r))

    def immersed_boundary_diagnostics(self):
        t = np.array(self.ib_total_force, dtype=np.float64)
        return {
            "ib_active_cell_count": int(self.counters[5]),
            "ib_clipped_cell_count": int(self.counters[6]),
            "ib_total_force_x": float(t[0]),
            "ib_total_force_y": float(t[1]),
            "ib_total_force_z": float(t[2]),
            "ib_total_force_norm": float(np.linalg.norm(t)),
            "contact_candidate_count": int(self.counters[3]),
            "contact_damped_particle_count": int(self.counters[4]),
        }

    def coupling_diagnostics(self):
        w = self.particle_valid_weights_numpy()
        return {
            "unsupported_particle_count": int(self.counters[0]),
            "partial_support_particle_count": int(self.counters[1]),
            "clipped_particle_count": int(self.counters[2]),
            "min_particle_valid_weight": float(w.min()) if len(w) else 0.0,
            "mean_particle_valid_weight": float(w.mean()) if len(w) else 0.0,
            "effective_gamma": self.effective_gamma(),
            **self.immersed_boundary_diagnostics(),
        }


class OracleSimulation:
    """fsi/simulation.py:12-153 on the CPU (snapshots are not part of the oracle)."""

    def __init__(self, config, precision: str = "f32"):
        config.validate()
        self.config = config
        self.lbm = OracleLBM(config.lbm, precision)
        self.mpm = OracleMPM(config.mpm, precision)
        self.coupler = OracleCoupler(config.coupling, self.lbm, self.mpm)
        self.step_index = 0
        self.time = 0.0
        self._lbm_initialized = False
        self._mpm_initialized = False

    def initialize_lbm(self, solid_np=None):
        self.lbm.initialize(solid_np=solid_np)
        self._lbm_initialized = True

    def initialize_mpm_box(self, lower, upper, spacing=None, initial_velocity=(0.0, 0.0, 0.0)):
        n = self.mpm.initialize_particles_box(lower, upper, spacing, initial_velocity)
        self._mpm_initialized = True
        return n

    def initialize_mpm_from_numpy(self, positions, velocities=None, particle_mass=None,
                                  particle_volume=None):
        n = self.mpm.initialize_particles_from_numpy(positions, velocities, particle_mass,
                                                     particle_volume)
        self._mpm_initialized = True
        return n

    def initialize(self, solid_np=None, mpm_box_lower=None, mpm_box_upper=None,
                   mpm_box_spacing=None, mpm_initial_velocity=(0.0, 0.0, 0.0)):
        if mpm_box_lower is None or mpm_box_upper is None:
            raise ValueError("mpm_box_lower and mpm_box_upper are required.")
        self.initialize_lbm(solid_np)
        return self.initialize_mpm_box(mpm_box_lower, mpm_box_upper, mpm_box_spacing,
                                       mpm_initial_velocity)

    def _validate_initialized(self):
        if not self._lbm_initialized or not self.lbm._initialized:
            raise RuntimeError("initialize_lbm must be called before stepping.")
        if not self._mpm_initialized or not self.mpm._initialized:
            raise RuntimeError("MPM initialization must be called before stepping.")

    def step(self):
        self._validate_initialized()
        self.coupler.step(self.config.lbm_dt)
        self.step_index += 1
        self.time += self.config.lbm_dt
        return self.diagnostics()

    def run(self, steps=None, write_output=False, write_initial=False):
        self._validate_initialized()
        count = self.config.num_steps if steps is None else int(steps)
        if count <= 0:
            raise ValueError("steps must be positive.")
        return [self.step() for _ in range(count)]

    def diagnostics(self):
        cd = self.coupler.coupling_diagnostics()
        return {
            "step": self.step_index,
            "time": self.time,
            "lbm_total_mass": self.lbm.total_mass(),
            "lbm_max_velocity_norm": self.lbm.max_velocity_norm(),
            "mpm_particle_count": self.mpm.particle_count(),
            "mpm_center_of_mass": self.mpm.center_of_mass(),
            "mpm_max_velocity_norm": self.mpm.max_velocity_norm(),
            "total_particle_coupling_force": self.coupler.total_particle_coupling_force(),
            "total_fluid_coupling_force": self.coupler.total_fluid_coupling_force(),
            "coupling_unsupported_particle_count": cd["unsupported_particle_count"],
            "coupling_partial_support_particle_count": cd["partial_support_particle_count"],
            "coupling_clipped_particle_count": cd["clipped_particle_count"],
            "coupling_effective_gamma": cd["effective_gamma"],
            "coupling_ib_active_cell_count": cd["ib_active_cell_count"],
            "coupling_ib_clipped_cell_count": cd["ib_clipped_cell_count"],
            "coupling_ib_total_force_norm": cd["ib_total_force_norm"],
            "coupling_contact_candidate_count": cd["contact_candidate_count"],
            "coupling_contact_damped_particle_count": cd["contact_damped_particle_count"],
        }


def polar_rotation(F: np.ndarray, precision: str = "f64") -> np.ndarray:
    """Rotation factor of one 3x3 matrix via the oracle's SVD (test helper)."""
    dt = np.float32 if precision == "f32" else np.float64
    Fa = np.ascontiguousarray(F, dtype=dt)
    R = np.zeros((3, 3), dt)
    getattr(lib(), f"oref_polar_rotation_{precision}")(_p(Fa), _p(R))
    return R
