"""D3Q19 MRT lattice Boltzmann solver on the GPU (reference: fsi/lbm3d.py).

Same constructor, methods, accessor shapes/dtypes and exceptions as the
reference's ``LBMSolver3D`` (fsi/lbm3d.py:57-534). State lives in torch device
tensors; every update is a HIP kernel of ``libfsi_hip.so``.

Between steps the lattice is in one of two states:

* materialised - ``pop[cur]`` holds the reference's ``f`` (post-stream,
  post-boundary populations, 19 runs of nz values per lattice row) and
  ``rho``/``vel`` hold its macroscopic fields. ``initialize`` leaves this
  state.
* post-collision - ``pop[cur]`` holds the populations right after the last
  collision; streaming is deferred to the next kernel. ``step`` leaves this
  state.

A step from the post-collision state is one fused pull kernel (stream,
x-face boundaries, macroscopic fields, collision). Anything that needs the
reference's post-step state - the accessors, the coupler, and any change of
the force field (the macroscopic velocity must use the force of the step
that produced it) - first materialises it; the next step then only collides.
FP16 storage keeps (f - w rho0) * pop_scale; I16 storage keeps
(f - w rho0) / w * s as 16-bit integers, the kernels adapting s to the
largest stored value through a small device control block (the host only
counts the population writes).

When the host knows the force is uniform on fluid cells
(``_force_uniform``), kernels take the value as a constant and read no force
memory.
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import torch

from . import _native as nat
from .config import BoundaryConfig, LBMConfig

_Q = 19
_SUPPORTED_X_BOUNDARIES = ("periodic", "pressure", "velocity")
# FP16 storage scale: deviations up to 2.0 fit, the subnormal range starts
# below 2e-9
POP_SCALE_FP16 = 2.0 ** 15


def _boundary_mode(boundary: BoundaryConfig) -> int:
    # fsi/lbm3d.py:16-27
    if boundary.boundary_type == "wall":
        raise NotImplementedError(
            "wall face boundaries are reserved for a later step; use solid masks for now."
        )
    if boundary.boundary_type not in nat.BC_MODES:
        raise ValueError(f"Unsupported boundary type: {boundary.boundary_type}.")
    return nat.BC_MODES[boundary.boundary_type]


def relaxation_rates(viscosity: float) -> tuple[float, float]:
    """fsi/lbm3d.py:141-143: tau = 3 nu + 0.5, s_v = 1/tau, s_o = 8(2 - s_v)/(8 - s_v)."""
    tau = 3.0 * viscosity + 0.5
    s_v = 1.0 / tau
    s_o = 8.0 * (2.0 - s_v) / (8.0 - s_v)
    return s_v, s_o


class LBMSolver3D:
    """Dense single-phase D3Q19/MRT LBM solver running on one GPU."""

    def __init__(self, config: LBMConfig):
        self.config = config
        self.config.validate()
        self._validate_supported_config()

        self.nx = int(config.nx)
        self.ny = int(config.ny)
        self.nz = int(config.nz)
        self.q = _Q
        self.n_cells = self.nx * self.ny * self.nz
        self.bc_x_left_mode = _boundary_mode(config.x_left)
        self.bc_x_right_mode = _boundary_mode(config.x_right)
        self.storage = config.storage
        self.pop_scale = {"fp16": POP_SCALE_FP16, "i16": self._i16_initial_scale()}.get(self.storage, 1.0)

        self.device = nat.device()
        self.real_dtype = torch.float64 if self.storage == "fp64" else torch.float32
        self.real_code = nat.REAL_F64 if self.storage == "fp64" else nat.REAL_F32
        self._np_real = np.float64 if self.storage == "fp64" else np.float32
        pop_dtype = {"fp32": torch.float32, "fp16": torch.float16, "i16": torch.int16,
                     "fp64": torch.float64}[self.storage]

        n = self.n_cells
        dev, rd = self.device, self.real_dtype
        self.pop = [torch.empty(_Q * n, dtype=pop_dtype, device=dev) for _ in range(2)]
        self._cur = 0
        self.rho = torch.empty(n, dtype=rd, device=dev)
        self.vel = torch.empty(3 * n, dtype=rd, device=dev)
        self.force = torch.zeros(3 * n, dtype=rd, device=dev)
        self.solid = torch.zeros(n, dtype=torch.int32, device=dev)
        self.nbmask = torch.zeros(n, dtype=torch.int32, device=dev)
        # two slots (previous / next step) x two faces x 3 components
        self.face_vel = torch.zeros(12 * self.ny * self.nz, dtype=rd, device=dev)
        self._sum_ws, self._sum_out = nat.reduction_buffers(dev)
        # I16 storage: scale and range of the last population writes (device)
        self.store_ctrl = (torch.zeros(nat.STORE_CTRL_FLOATS, dtype=torch.float32, device=dev)
                           if self.storage == "i16" else None)
        self._epoch = 0  # population writes since initialize()

        # lattices whose z-rows are whole wavefronts read the solid flags as
        # a row bitmap (scalar loads per wavefront) instead of the 4-byte
        # neighbour mask per cell; set to None before initialize() to keep
        # the mask
        self.solid_bits = (torch.zeros(nat.solid_bits_words(self.nx, self.ny, self.nz), dtype=torch.int32,
                                       device=dev)
                           if self.nz % 64 == 0 else None)

        self._params = self._make_params()
        self._buffers = self._make_buffers()
        self._materialized = False
        self._face_slot = 0
        self._force_dirty = False
        self._force_uniform: tuple[float, float, float] | None = None
        self._initialized = False

    # ------------------------------------------------------------ setup --

    def _validate_supported_config(self) -> None:
        # fsi/lbm3d.py:85-108
        c = self.config
        pop_bytes = {"fp32": 4, "fp16": 2, "i16": 2, "fp64": 8}.get(getattr(c, "storage", "fp32"), 4)
        if _Q * int(c.nx) * int(c.ny) * int(c.nz) * pop_bytes >= 2 ** 32:
            # the lattice kernels address a population buffer with 32-bit
            # byte offsets (one VGPR per address, SGPR base)
            raise NotImplementedError(
                f"a {c.nx}x{c.ny}x{c.nz} lattice needs {_Q * pop_bytes} B x cells >= 4 GiB per population "
                f"buffer in {getattr(c, 'storage', 'fp32')} storage; the lattice kernels support up to "
                f"{(2 ** 32 - 1) // (_Q * pop_bytes):,} cells (use fp16 storage or a smaller lattice)."
            )
        if self.config.use_sparse:
            raise NotImplementedError("Sparse LBM storage is not implemented in Step 2.")
        for name, boundary in (("x_left", self.config.x_left), ("x_right", self.config.x_right)):
            if boundary.boundary_type not in _SUPPORTED_X_BOUNDARIES:
                raise NotImplementedError(
                    f"{name}={boundary.boundary_type!r} is not implemented in Step 2."
                )
        for name, boundary in (
            ("y_left", self.config.y_left),
            ("y_right", self.config.y_right),
            ("z_left", self.config.z_left),
            ("z_right", self.config.z_right),
        ):
            if boundary.boundary_type != "periodic":
                raise NotImplementedError(
                    f"{name}={boundary.boundary_type!r} is not implemented in Step 2; "
                    "use periodic Y/Z boundaries and solid masks for now."
                )

    def _make_params(self) -> nat.LbmParams:
        c = self.config
        p = nat.LbmParams()
        p.nx, p.ny, p.nz = self.nx, self.ny, self.nz
        p.storage = nat.STORAGE_CODES[self.storage]
        p.rho0 = float(c.rho0)
        p.initial_velocity = nat.c_double3(*[float(v) for v in c.initial_velocity])
        p.s_v, p.s_o = relaxation_rates(float(c.viscosity))
        p.left_mode = self.bc_x_left_mode
        p.right_mode = self.bc_x_right_mode
        p.bc_rho = (nat.ctypes.c_double * 2)(float(c.x_left.rho), float(c.x_right.rho))
        p.bc_vel[0] = nat.c_double3(*[float(v) for v in c.x_left.velocity])
        p.bc_vel[1] = nat.c_double3(*[float(v) for v in c.x_right.velocity])
        p.pop_scale = float(self.pop_scale)
        p.store_epoch = 0
        return p

    def _i16_initial_scale(self) -> float:
        """Scale of the I16 populations fsi_lbm_init writes: the largest
        normalised deviation (f - w rho0)/w = (rho - rho0) + rho (3 e.u +
        4.5 (e.u)^2 - 1.5 u.u) the configured velocities, face densities and
        force give (|e.u| <= sqrt(2)|u|), at a quarter of the 16-bit range,
        and no finer than for a deviation of 2.5e-3 (a flow the coupling
        starts); the lattice kernels adapt it within the first steps."""
        c = self.config
        speeds = [float(np.linalg.norm(c.initial_velocity))]
        drho = 0.0
        for face in (c.x_left, c.x_right):
            if face.boundary_type == "velocity":
                speeds.append(float(np.linalg.norm(face.velocity)))
            elif face.boundary_type == "pressure":
                drho = max(drho, abs(float(face.rho) - float(c.rho0)))
        v = max(speeds)
        est = drho + (1.0 + drho) * (3.0 * np.sqrt(2.0) * v + 10.5 * v * v) + 3.0 * float(np.linalg.norm(c.force))
        return 0.25 / max(est, 2.5e-3)

    @staticmethod
    def _i16_tracked(epoch: int) -> bool:
        # the writes whose kernels record their largest value (lbm.hip)
        return epoch < 64 or epoch % 4 == 0

    def storage_range(self) -> float:
        """I16 storage: the largest |stored value| of the last tracked
        population write (every one of the first 64 writes, then every
        fourth) as a fraction of the 16-bit range: above 1, values were
        clipped; 0 right after initialize. 0 for the other storages."""
        if self.store_ctrl is None:
            return 0.0
        t = self._epoch
        while t > 0 and not self._i16_tracked(t):
            t -= 1
        slot = (t % 3) * nat.STORE_CTRL_SLOT
        parts = self.store_ctrl[slot:slot + nat.STORE_CTRL_SLOT].view(-1, nat.STORE_CTRL_LINE)[1:, 0]
        return float(parts.max().item())

    def storage_scale(self) -> float:
        """Scale s of the current I16 populations ((f - w rho0)/w * s is
        stored as round(32767 x)); the fixed FP16 scale, 1 otherwise."""
        if self.store_ctrl is None:
            return float(self.pop_scale)
        return float(self.store_ctrl[(self._epoch % 3) * nat.STORE_CTRL_SLOT].item())

    def initialize(self, solid_np: np.ndarray | None = None) -> None:
        # fsi/lbm3d.py:193-203
        if solid_np is None:
            solid = np.zeros((self.nx, self.ny, self.nz), dtype=np.int32)
        else:
            solid = self._normalize_solid_mask(solid_np)
        self.solid.copy_(torch.from_numpy(solid.reshape(-1)))
        self._buffers = self._make_buffers()
        if self.store_ctrl is not None:
            nat.fill_zero(self.store_ctrl)
        self._epoch = 0
        self._params.store_epoch = 0
        nat.call("fsi_lbm_init", nat.ref(self._params), nat.ref(self._buffers),
                 nat.ptr(self.pop[0]), nat.ptr(self.pop[1]))
        self._cur = 0
        self._face_slot = 0
        self._materialized = True
        self._force_dirty = False
        self._initialized = True
        self.reset_force()

    def _make_buffers(self) -> nat.LbmBuffers:
        return nat.LbmBuffers(
            nat.ptr(self.solid), nat.ptr(self.nbmask), nat.ptr(self.rho), nat.ptr(self.vel),
            nat.ptr(self.force), nat.ptr(self.face_vel), nat.ptr(self.store_ctrl), nat.ptr(self.solid_bits),
        )

    def _normalize_solid_mask(self, solid_np: np.ndarray) -> np.ndarray:
        arr = np.asarray(solid_np)
        expected_shape = (self.nx, self.ny, self.nz)
        if arr.shape != expected_shape:
            raise ValueError(f"solid_np shape must be {expected_shape}, got {arr.shape}.")
        return np.ascontiguousarray(np.where(arr > 0, 1, 0).astype(np.int32))

    @classmethod
    def load_solid_from_txt(cls, filename: str | Path, nx: int, ny: int, nz: int) -> np.ndarray:
        # fsi/lbm3d.py:213-217: Fortran-order text mask
        data = np.loadtxt(filename)
        data = np.reshape(data, (nx, ny, nz), order="F")
        return np.ascontiguousarray(np.where(data > 0, 1, 0).astype(np.int32))

    # ------------------------------------------------------------- step --

    @property
    def populations(self) -> torch.Tensor:
        """Current population buffer (storage dtype; population q of cell
        (i, j, k) at ((i*ny + j)*19 + q)*nz + k)."""
        return self.pop[self._cur]

    def _force_arg(self) -> nat.LbmForce:
        f = nat.LbmForce()
        if self._force_uniform is not None:
            f.field = None
            f.uniform = nat.c_double3(*self._force_uniform)
        else:
            f.field = nat.ptr(self.force).value
            f.uniform = nat.c_double3(0.0, 0.0, 0.0)
        return f

    # I16 storage: the first population writes after initialize() are
    # measured first (a pass without stores), so their scale fits what they
    # store whatever force or boundary starts the flow
    _MEASURED_WRITES = 3

    def _launch(self, phase: int, assembly=None) -> None:
        self._params.store_epoch = self._epoch + 1
        pin, pout = nat.ptr(self.pop[self._cur]), nat.ptr(self.pop[1 - self._cur])
        if assembly is None:
            name = "fsi_lbm_step"
            f = self._force_arg()
            args = (nat.ref(self._params), nat.ref(self._buffers), pin, pout, nat.ref(f), phase, self._face_slot)
        else:  # the collide phase with the force assembled in the same pass
            name = "fsi_lbm_collide_assembled"
            a, cforce, ib_force, ib_total, counters = assembly
            args = (nat.ref(self._params), nat.ref(self._buffers), nat.ref(a), pin, pout, nat.ptr(cforce),
                    nat.ptr(ib_force), nat.ptr(ib_total), nat.ptr(counters))
        if self.store_ctrl is not None and self._epoch < self._MEASURED_WRITES:
            m = 3 * nat.STORE_CTRL_SLOT
            nat.fill_zero(self.store_ctrl[m:m + nat.STORE_CTRL_SLOT])
            self._params.store_pass = 1
            nat.call(name, *args)
            self._params.store_pass = 2
        nat.call(name, *args)
        self._params.store_pass = 0
        self._epoch += 1
        self._cur = 1 - self._cur
        if phase != nat.PHASE_COLLIDE and self._faces:
            self._face_slot = 1 - self._face_slot

    @property
    def _faces(self) -> bool:
        return self.bc_x_left_mode != 0 or self.bc_x_right_mode != 0

    def step(self) -> None:
        # fsi/lbm3d.py:427-433
        if not self._initialized:
            raise RuntimeError("LBMSolver3D.initialize must be called before step().")
        self._launch(nat.PHASE_COLLIDE if self._materialized else nat.PHASE_STREAM_COLLIDE)
        self._materialized = False
        self._force_dirty = False

    def collide_assembled(self, assembly, cforce, ib_force, ib_total, counters) -> None:
        """The coupled step's force assembly (fsi/coupling.py:162-170: phi
        clamp, reset_force, coupling force, IB drag; ``assembly`` a
        ForceAssembly) fused with the collision of the next lattice step
        (lbm3d.py:314-333), which uses that force: one pass from the
        materialised state. The lattice force field holds the assembled
        force afterwards, as after fsi_force_assemble + step()."""
        if not self._initialized:
            raise RuntimeError("LBMSolver3D.initialize must be called before step().")
        self._ensure_macro()
        self._launch(nat.PHASE_COLLIDE, (assembly, cforce, ib_force, ib_total, counters))
        self._materialized = False
        self._force_dirty = False
        self._force_uniform = None

    def _ensure_macro(self) -> None:
        """Materialise the reference's post-step state: f, rho, u (lbm3d.py:409-425)."""
        if self._materialized:
            return
        if self._force_dirty:  # cannot happen: force updates materialise first
            raise RuntimeError("internal: post-collision state with a changed force")
        self._launch(nat.PHASE_STREAM_MACRO)
        self._materialized = True

    def _before_force_update(self) -> None:
        if self._initialized:
            self._ensure_macro()
        self._force_dirty = True

    # -------------------------------------------------------- force API --

    def _force_op(self, op: int, value=(0.0, 0.0, 0.0)) -> None:
        self._before_force_update()
        v = (nat.ctypes.c_double * 3)(*[float(x) for x in value])
        nat.call("fsi_lbm_force_op", nat.ref(self._params), nat.ref(self._buffers), op, v,
                 nat.ptr(self.force), nat.ptr(self.force))

    def _rounded(self, value) -> tuple[float, float, float]:
        arr = np.asarray(value, dtype=self._np_real)
        return float(arr[0]), float(arr[1]), float(arr[2])

    def clear_force(self) -> None:
        self._force_op(0)
        self._force_uniform = (0.0, 0.0, 0.0)

    def reset_force(self) -> None:
        # fsi/lbm3d.py:505-510: fluid = LBMConfig.force, solid = 0
        value = self._rounded(self.config.force)
        self._force_op(1, value)
        self._force_uniform = value

    def set_uniform_force(self, force: tuple[float, float, float]) -> None:
        value = self._rounded(self._validate_force_vector(force, "force"))
        self._force_op(1, value)
        self._force_uniform = value

    def add_uniform_force(self, force: tuple[float, float, float]) -> None:
        value = self._rounded(self._validate_force_vector(force, "force"))
        self._force_op(2, value)
        if self._force_uniform is not None:
            old = np.asarray(self._force_uniform, dtype=self._np_real)
            self._force_uniform = tuple(float(x) for x in old + np.asarray(value, self._np_real))

    def set_force_from_numpy(self, force_np: np.ndarray) -> None:
        force = self._normalize_force_field(force_np)
        self._before_force_update()
        staged = torch.from_numpy(force.reshape(-1)).to(self.device)
        nat.call("fsi_lbm_force_import", nat.ref(self._params), nat.ref(self._buffers),
                 nat.ptr(staged))
        self._force_uniform = None

    def zero_force_on_solid(self) -> None:
        self._force_op(3)

    def set_force_field(self, force: torch.Tensor, uniform: tuple | None = None) -> None:
        """Device-side replacement of the whole force field (used by the coupler)."""
        self._before_force_update()
        if force.data_ptr() != self.force.data_ptr():
            self.force.copy_(force)
        self._force_uniform = uniform

    def total_force(self) -> np.ndarray:
        force_np = self.force_numpy()
        solid_np = self.solid_numpy()
        return force_np[solid_np == 0].sum(axis=0)

    @staticmethod
    def _validate_force_vector(force, name: str) -> tuple[float, float, float]:
        arr = np.asarray(force, dtype=np.float32)
        if arr.shape != (3,):
            raise ValueError(f"{name} must have 3 components.")
        return float(arr[0]), float(arr[1]), float(arr[2])

    def _normalize_force_field(self, force_np: np.ndarray) -> np.ndarray:
        arr = np.asarray(force_np, dtype=np.float32)
        expected_shape = (self.nx, self.ny, self.nz, 3)
        if arr.shape != expected_shape:
            raise ValueError(f"force_np shape must be {expected_shape}, got {arr.shape}.")
        return np.ascontiguousarray(arr)

    # -------------------------------------------------------- accessors --

    def _aos_device(self, soa: torch.Tensor, ncomp: int) -> torch.Tensor:
        out = torch.empty(soa.numel(), dtype=soa.dtype, device=self.device)
        nat.call("fsi_soa_to_aos", self.real_code, self.n_cells, ncomp, nat.ptr(soa), nat.ptr(out))
        return out.view(self.nx, self.ny, self.nz, ncomp)

    # Device accessors: the *_numpy arrays' contents, shapes and dtypes as
    # device tensors (fresh tensors or views of the state: copy before the
    # next step changes them); the *_numpy accessors copy them to the host.

    def density_device(self) -> torch.Tensor:
        self._ensure_macro()
        return self.rho.view(self.nx, self.ny, self.nz)

    def velocity_device(self) -> torch.Tensor:
        self._ensure_macro()
        return self._aos_device(self.vel, 3)

    def force_device(self) -> torch.Tensor:
        return self._aos_device(self.force, 3)

    def solid_device(self) -> torch.Tensor:
        return self.solid.view(self.nx, self.ny, self.nz)

    def density_numpy(self) -> np.ndarray:
        return self.density_device().cpu().numpy()

    def velocity_numpy(self) -> np.ndarray:
        return self.velocity_device().cpu().numpy()

    def solid_numpy(self) -> np.ndarray:
        return self.solid_device().cpu().numpy()

    def distribution_numpy(self) -> np.ndarray:
        self._ensure_macro()
        out = torch.empty(self.n_cells * _Q, dtype=self.real_dtype, device=self.device)
        self._params.store_epoch = self._epoch
        nat.call("fsi_lbm_export_populations", nat.ref(self._params), nat.ref(self._buffers),
                 nat.ptr(self.pop[self._cur]),
                 nat.ptr(out))
        return out.cpu().numpy().reshape(self.nx, self.ny, self.nz, _Q)

    def force_numpy(self) -> np.ndarray:
        return self.force_device().cpu().numpy()

    def lattice_sums(self, cforce: torch.Tensor | None = None) -> torch.Tensor:
        """FP64 sums over the materialised state, on the device (5 doubles):
        fluid mass, max |u|^2, and the coupling buffer's force summed over
        fluid cells (0 without one). No host copy."""
        self._ensure_macro()
        nat.call("fsi_reduce_lattice", nat.ref(self._params), nat.ref(self._buffers), nat.ptr(cforce),
                 nat.ptr(self._sum_ws), self._sum_ws.numel(), nat.ptr(self._sum_out))
        return self._sum_out[:5]

    # The accessors below keep the reference's float32 host reductions (its
    # validation values depend on their rounding); the per-step diagnostics
    # use lattice_sums().

    def total_mass(self) -> float:
        # fsi/lbm3d.py:449-452
        rho_np = self.density_numpy()
        solid_np = self.solid_numpy()
        return float(rho_np[solid_np == 0].sum())

    def max_velocity_norm(self) -> float:
        # fsi/lbm3d.py:454-456
        velocity_np = self.velocity_numpy()
        return float(np.linalg.norm(velocity_np, axis=-1).max())

    # device views used by the coupler
    def macro_tensors(self) -> tuple[torch.Tensor, torch.Tensor]:
        self._ensure_macro()
        return self.rho, self.vel
