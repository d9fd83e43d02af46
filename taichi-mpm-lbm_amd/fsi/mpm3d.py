"""MLS-MPM elastic solid on the GPU (reference: fsi/mpm3d.py).

Same constructor, methods, accessor shapes/dtypes and exceptions as the
reference's ``MPMSolver3D`` (fsi/mpm3d.py:9-368). Particle state is SoA in
device tensors of ``max_particles`` entries per component; the background
grid is one AoS (momentum/velocity xyz, mass) tensor per node. A substep is
three launches: grid clear, P2G, grid update + G2P.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _native as nat
from .config import MPMConfig


class MPMSolver3D:
    """Standalone dense 3D elastic MLS-MPM solver running on one GPU.

    Coordinates are grid/lattice coordinates, as in the reference.
    """

    def __init__(self, config: MPMConfig, precision: str = "fp32"):
        self.config = config
        self.config.validate()
        if precision not in ("fp32", "fp64"):
            raise ValueError(f"Unsupported MPM precision: {precision}.")

        self.nx = int(config.nx)
        self.ny = int(config.ny)
        self.nz = int(config.nz)
        self.max_particles = int(config.max_particles)
        self.dx = float(config.dx)
        self.inv_dx = 1.0 / self.dx
        self.dt_default = float(config.dt)
        self.boundary_width = int(config.boundary_width)
        self.boundary_damping = float(config.boundary_damping)
        self.apic = bool(config.apic)
        # fsi/mpm3d.py:33-38
        self.mu = config.youngs_modulus / (2.0 * (1.0 + config.poisson_ratio))
        self.lam = (
            config.youngs_modulus
            * config.poisson_ratio
            / ((1.0 + config.poisson_ratio) * (1.0 - 2.0 * config.poisson_ratio))
        )

        self.precision = precision
        self.device = nat.device()
        self.real_dtype = torch.float64 if precision == "fp64" else torch.float32
        self.real_code = nat.REAL_F64 if precision == "fp64" else nat.REAL_F32
        self._np_real = np.float64 if precision == "fp64" else np.float32
        cap, dev, rd = self.max_particles, self.device, self.real_dtype
        self.x = torch.zeros(3 * cap, dtype=rd, device=dev)
        self.v = torch.zeros(3 * cap, dtype=rd, device=dev)
        self.C = torch.zeros(9 * cap, dtype=rd, device=dev)
        self.F = torch.zeros(9 * cap, dtype=rd, device=dev)
        self.mass = torch.zeros(cap, dtype=rd, device=dev)
        self.volume = torch.zeros(cap, dtype=rd, device=dev)
        self.active = torch.zeros(cap, dtype=torch.int32, device=dev)
        self.particle_force = torch.zeros(3 * cap, dtype=rd, device=dev)
        self.grid = torch.zeros(4 * self.nx * self.ny * self.nz, dtype=rd, device=dev)
        # second grid: batched substeps alternate the two, each G2P pass
        # zeroing the one the next P2G pass fills
        self.grid_alt = torch.zeros_like(self.grid)
        # third grid: batched substeps then run G2P of one substep and P2G of
        # the next in one launch, rotating the three grids
        self.grid_alt2 = torch.zeros_like(self.grid)
        # node boxes each batched P2G pass writes (8 int32 per workgroup,
        # three passes), so the pass clearing that grid zeroes only those
        # boxes; one entry per possible workgroup: sort chunks or unsorted
        # blocks
        self._zero_boxes_cap = self._max_workgroups()
        self._zero_boxes = torch.zeros(16 + 24 * self._zero_boxes_cap, dtype=torch.int32, device=dev)
        self._sum_ws, self._sum_out = nat.reduction_buffers(dev)
        # particle order: slot t holds input particle ids[t]; sort_particles()
        # reorders the slots by cell key (buffers for it are made on first use)
        self.ids = torch.arange(cap, dtype=torch.int32, device=dev)
        self._sort = None
        self._permuted = False
        self._num_chunks = 0  # P2G work list length after a sort (0: slot order)
        # FP32 substeps with a work list of at most this many chunks run three
        # wavefronts per particle (one per x plane of the stencil): a small
        # particle set is bound by one wavefront's latency, not by throughput
        # (measured on MI355X: 0.66x the time at 33-51 chunks, 0.83x at 112,
        # 1.23x at 168)
        self.split_max_chunks = 128
        # particle passes of at most this many workgroups (2 per CU on the
        # 256 CUs) gather the 27 stencil nodes in one batch; larger ones a
        # plane of 9 at a time with fewer registers (0: always the latter)
        self.wide_max_workgroups = 512
        self.num_particles = 0
        # optional per-particle Lame parameters (2 mu, lambda) in slot order,
        # 2 x cap (set_particle_material); None: the config's E and nu
        self.material = None
        self._buffers = self._make_buffers()
        self._initialized = False

    def _max_workgroups(self) -> int:
        """Upper bound on P2G workgroups: sort chunks (at most cap/256 plus
        one per occupied cell block plus one) or ceil(cap/256) unsorted."""
        cap = self.max_particles
        blocks = ((self.nx + 5) // 4) * ((self.ny + 5) // 4) * ((self.nz + 5) // 4) + 1
        return cap // 256 + min(blocks, cap) + 1

    # ----------------------------------------------------------- set-up --

    def initialize_particles_box(
        self,
        lower: tuple[float, float, float],
        upper: tuple[float, float, float],
        spacing: float | tuple[float, float, float] | None = None,
        initial_velocity: tuple[float, float, float] = (0.0, 0.0, 0.0),
    ) -> int:
        # fsi/mpm3d.py:69-96
        lower_np = self._validate_vec3(lower, "lower")
        upper_np = self._validate_vec3(upper, "upper")
        if np.any(upper_np <= lower_np):
            raise ValueError("upper must be greater than lower in every dimension.")
        spacing_np = self._default_spacing() if spacing is None else self._validate_spacing(spacing)
        axes = [np.arange(lower_np[a], upper_np[a], spacing_np[a]) for a in range(3)]
        if any(len(axis_values) == 0 for axis_values in axes):
            raise ValueError("particle box produced no particles.")
        mesh = np.meshgrid(*axes, indexing="ij")
        positions = np.stack([component.ravel() for component in mesh], axis=1).astype(np.float32)
        particle_volume = float(np.prod(spacing_np))
        velocities = np.broadcast_to(
            self._validate_vec3(initial_velocity, "initial_velocity"), positions.shape
        ).astype(np.float32)
        return self.initialize_particles_from_numpy(
            positions, velocities=velocities, particle_volume=particle_volume
        )

    def initialize_particles_from_numpy(
        self,
        positions: np.ndarray,
        velocities: np.ndarray | None = None,
        particle_mass: float | None = None,
        particle_volume: float | None = None,
    ) -> int:
        # fsi/mpm3d.py:98-165
        pos = np.asarray(positions, dtype=np.float32)
        if pos.ndim != 2 or pos.shape[1] != 3:
            raise ValueError("positions must have shape (n, 3).")
        count = int(pos.shape[0])
        if count <= 0:
            raise ValueError("positions must contain at least one particle.")
        if count > self.max_particles:
            raise ValueError(f"particle count {count} exceeds max_particles={self.max_particles}.")
        if velocities is None:
            vel = np.zeros((count, 3), dtype=np.float32)
        else:
            vel = np.asarray(velocities, dtype=np.float32)
            if vel.shape != (count, 3):
                raise ValueError(f"velocities must have shape {(count, 3)}, got {vel.shape}.")
        volume_value = (
            float(particle_volume)
            if particle_volume is not None
            else self.dx ** 3 / float(self.config.particles_per_cell ** 3)
        )
        if volume_value <= 0.0:
            raise ValueError("particle_volume must be positive.")
        mass_value = (
            float(particle_mass) if particle_mass is not None else float(self.config.density) * volume_value
        )
        if mass_value <= 0.0:
            raise ValueError("particle_mass must be positive.")

        cap = self.max_particles
        rd = self._np_real
        # host staging in SoA planes; host values pass through float32 first
        # because the reference's fields are float32
        x = np.zeros((3, cap), rd)
        v = np.zeros((3, cap), rd)
        F = np.zeros((9, cap), rd)
        x[:, :count] = pos.T
        v[:, :count] = vel.T
        for d in (0, 4, 8):
            F[d, :count] = 1.0
        mass = np.zeros(cap, rd)
        vol = np.zeros(cap, rd)
        mass[:count] = np.float32(mass_value)
        vol[:count] = np.float32(volume_value)
        active = np.zeros(cap, np.int32)
        active[:count] = 1
        self.x.copy_(torch.from_numpy(x.reshape(-1)))
        self.v.copy_(torch.from_numpy(v.reshape(-1)))
        self.F.copy_(torch.from_numpy(F.reshape(-1)))
        nat.fill_zero(self.C)
        self.mass.copy_(torch.from_numpy(mass))
        self.volume.copy_(torch.from_numpy(vol))
        self.active.copy_(torch.from_numpy(active))
        nat.fill_zero(self.particle_force)
        self.ids.copy_(torch.arange(cap, dtype=torch.int32))
        self._permuted = False
        self._num_chunks = 0
        self.num_particles = count
        self.material = None
        self._buffers = self._make_buffers()
        self._initialized = True
        return count

    def set_particle_material(self, youngs_modulus, poisson_ratio=None) -> None:
        """Per-particle elastic material (no counterpart in the reference,
        whose MPMConfig has one global E and nu, fsi/mpm3d.py:33-38): Young's
        modulus and Poisson ratio per particle in input order (scalars
        broadcast; poisson_ratio defaults to the config's). The P2G stress
        then uses each particle's own Lame parameters, computed as the
        reference computes the global ones. initialize_particles_* clears it."""
        if not self._initialized:
            raise RuntimeError("MPMSolver3D has no particles to give a material.")
        n = self.num_particles
        E = np.broadcast_to(np.asarray(youngs_modulus, dtype=np.float64), (n,))
        nu = np.broadcast_to(np.asarray(self.config.poisson_ratio if poisson_ratio is None else poisson_ratio,
                                        dtype=np.float64), (n,))
        if not (np.all(np.isfinite(E)) and np.all(E > 0.0)):
            raise ValueError("youngs_modulus must be positive and finite.")
        if not (np.all(nu > -1.0) and np.all(nu < 0.5)):
            raise ValueError("poisson_ratio must lie in (-1, 0.5).")
        mu = E / (2.0 * (1.0 + nu))
        lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
        cap = self.max_particles
        planes = np.zeros((2, cap), self._np_real)
        order = self.ids[:n].cpu().numpy() if self._permuted else np.arange(n)
        planes[0, :n] = 2.0 * mu[order]
        planes[1, :n] = lam[order]
        if self.material is None:
            self.material = torch.empty(2 * cap, dtype=self.real_dtype, device=self.device)
        self.material.copy_(torch.from_numpy(planes.reshape(-1)))
        if self._sort is not None and "material" not in self._sort["alt"]:
            self._sort["alt"]["material"] = torch.empty_like(self.material)
        self._buffers = self._make_buffers()

    def _default_spacing(self) -> np.ndarray:
        return np.full(3, self.dx / float(self.config.particles_per_cell), dtype=np.float32)

    @staticmethod
    def _validate_vec3(values, name: str) -> np.ndarray:
        arr = np.asarray(values, dtype=np.float32)
        if arr.shape != (3,):
            raise ValueError(f"{name} must have 3 components.")
        return arr

    @staticmethod
    def _validate_spacing(spacing) -> np.ndarray:
        if np.isscalar(spacing):
            arr = np.full(3, float(spacing), dtype=np.float32)
        else:
            arr = np.asarray(spacing, dtype=np.float32)
        if arr.shape != (3,):
            raise ValueError("spacing must be a scalar or have 3 components.")
        if np.any(arr <= 0.0):
            raise ValueError("spacing must be positive.")
        return arr

    # ---------------------------------------------------------- stepping --

    def params(self, dt: float) -> nat.MpmParams:
        c = self.config
        p = nat.MpmParams()
        p.nx, p.ny, p.nz = self.nx, self.ny, self.nz
        p.real = self.real_code
        p.num_particles = self.num_particles
        p.capacity = self.max_particles
        p.dx = self.dx
        p.inv_dx = self.inv_dx
        p.dt = float(dt)
        p.two_mu = 2.0 * self.mu
        p.lam = self.lam
        p.gravity = nat.c_double3(*[float(g) for g in c.gravity])
        p.boundary_width = self.boundary_width
        p.boundary_damping = self.boundary_damping
        p.apic = 1 if self.apic else 0
        p.num_chunks = self._num_chunks if self._permuted else 0
        p.split_max_chunks = self.split_max_chunks
        p.wide_max_workgroups = self.wide_max_workgroups
        p.clamp_lower = 0.5 * self.dx
        p.clamp_upper = nat.c_double3((self.nx - 1.5) * self.dx, (self.ny - 1.5) * self.dx,
                                      (self.nz - 1.5) * self.dx)
        return p

    def _dt_value(self, dt: float | None) -> float:
        dt_value = self.dt_default if dt is None else float(dt)
        if dt_value <= 0.0:
            raise ValueError("dt must be positive.")
        # the reference stores dt in a float32 field
        return float(self._np_real(dt_value))

    def substep(self, dt: float | None = None) -> None:
        # fsi/mpm3d.py:189-203
        if not self._initialized:
            raise RuntimeError(
                "MPMSolver3D.initialize_particles_box or initialize_particles_from_numpy "
                "must be called before substep()."
            )
        p = self.params(self._dt_value(dt))
        nat.call("fsi_mpm_clear_grid", nat.ref(p), nat.ref(self._buffers))
        nat.call("fsi_mpm_p2g", nat.ref(p), nat.ref(self._buffers), None, None, 0, 1)
        nat.call("fsi_mpm_grid_g2p", nat.ref(p), nat.ref(self._buffers), None, None)

    def _sort_state(self):
        if self._sort is None:
            cap, dev, rd = self.max_particles, self.device, self.real_dtype
            alt = {
                "x": torch.empty(3 * cap, dtype=rd, device=dev),
                "v": torch.empty(3 * cap, dtype=rd, device=dev),
                "C": torch.empty(9 * cap, dtype=rd, device=dev),
                "F": torch.empty(9 * cap, dtype=rd, device=dev),
                "mass": torch.empty(cap, dtype=rd, device=dev),
                "volume": torch.empty(cap, dtype=rd, device=dev),
                "active": torch.zeros(cap, dtype=torch.int32, device=dev),
                "particle_force": torch.empty(3 * cap, dtype=rd, device=dev),
                "ids": torch.empty(cap, dtype=torch.int32, device=dev),
            }
            if self.material is not None:
                alt["material"] = torch.empty_like(self.material)
            nbytes = ctypes.c_int64(0)
            p = self.params(self.dt_default)
            rc = nat.load_library().fsi_mpm_sort_workspace(nat.ref(p), ctypes.byref(nbytes))
            if rc != 0:
                raise RuntimeError(nat.load_library().fsi_last_error().decode(errors="replace"))
            ws = torch.empty(max(int(nbytes.value), 1), dtype=torch.uint8, device=dev)
            # P2G work list: at most cap/256 + (cell blocks + 1) chunks
            max_chunks = self._max_workgroups()
            chunks = torch.zeros(2 * max_chunks, dtype=torch.int32, device=dev)
            count = torch.zeros(1, dtype=torch.int32, device=dev)
            self._sort = {"alt": alt, "ws": ws, "chunks": chunks, "max_chunks": max_chunks, "count": count}
        return self._sort

    _SORTED_FIELDS = ("x", "v", "C", "F", "mass", "volume", "active", "particle_force")

    def _make_buffers(self) -> nat.MpmBuffers:
        chunks = self._sort["chunks"] if self._sort is not None else None
        return nat.MpmBuffers(
            nat.ptr(self.x), nat.ptr(self.v), nat.ptr(self.C), nat.ptr(self.F),
            nat.ptr(self.mass), nat.ptr(self.volume), nat.ptr(self.active),
            nat.ptr(self.particle_force), nat.ptr(self.grid), nat.ptr(self.grid_alt), nat.ptr(chunks),
            nat.ptr(self._zero_boxes), self._zero_boxes_cap, nat.ptr(self.grid_alt2), nat.ptr(self.material),
        )

    def sort_particles(self, keys_out: torch.Tensor | None = None,
                       perm_out: torch.Tensor | None = None) -> None:
        """Reorder the particle slots by cell key (a stable counting sort on
        the device; see fsi_mpm_sort). Results do not depend on the order
        beyond float rounding of the grid sums; the accessors return the
        particles in input order whatever the slot order."""
        if not self._initialized:
            raise RuntimeError("MPMSolver3D has no particles to sort.")
        st = self._sort_state()
        alt = st["alt"]
        dst = nat.MpmBuffers(
            nat.ptr(alt["x"]), nat.ptr(alt["v"]), nat.ptr(alt["C"]), nat.ptr(alt["F"]),
            nat.ptr(alt["mass"]), nat.ptr(alt["volume"]), nat.ptr(alt["active"]),
            nat.ptr(alt["particle_force"]), nat.ptr(self.grid), nat.ptr(self.grid_alt),
        )
        if self.material is not None:
            dst.material = nat.ptr(alt["material"])
        p = self.params(self.dt_default)
        nat.call("fsi_mpm_sort", nat.ref(p), nat.ref(self._buffers), nat.ref(dst), nat.ptr(self.ids),
                 nat.ptr(alt["ids"]), nat.ptr(st["ws"]), st["ws"].numel(), nat.ptr(keys_out),
                 nat.ptr(perm_out), nat.ptr(st["chunks"]), st["max_chunks"], nat.ptr(st["count"]))
        for name in self._SORTED_FIELDS + ("ids",) + (("material",) if self.material is not None else ()):
            cur = getattr(self, name)
            setattr(self, name, alt[name])
            alt[name] = cur
        self._buffers = self._make_buffers()
        self._permuted = True
        # the launch size of P2G needs the work list length on the host
        self._num_chunks = int(st["count"].item())

    def _input_order(self, arr: np.ndarray) -> np.ndarray:
        """Rows of a slot-ordered host array (first particle_count() slots)
        put back in input order."""
        if not self._permuted:
            return arr
        ids = self.ids[: self.num_particles].cpu().numpy()
        out = np.empty_like(arr)
        out[ids] = arr
        return out

    def clear_particle_force(self) -> None:
        # fsi/mpm3d.py:211-214
        nat.fill_zero(self.particle_force)

    # --------------------------------------------------------- accessors --

    def input_order_device(self, rows: torch.Tensor) -> torch.Tensor:
        """Rows of a slot-ordered device tensor (first particle_count()
        slots) in input order, as a fresh contiguous tensor."""
        if not self._permuted:
            return rows.contiguous()
        out = torch.empty(rows.shape, dtype=rows.dtype, device=rows.device)
        out.index_copy_(0, self.ids[: self.num_particles].long(), rows)
        return out

    def _plane_device(self, t: torch.Tensor, ncomp: int) -> torch.Tensor:
        n = self.num_particles
        return self.input_order_device(t.view(ncomp, self.max_particles)[:, :n].t())

    def particle_count(self) -> int:
        return int(self.num_particles)

    # Device accessors: the *_numpy arrays as fresh device tensors (input
    # order, same shapes and dtypes); the *_numpy accessors copy them.

    def positions_device(self) -> torch.Tensor:
        return self._plane_device(self.x, 3)

    def velocities_device(self) -> torch.Tensor:
        return self._plane_device(self.v, 3)

    def particle_forces_device(self) -> torch.Tensor:
        return self._plane_device(self.particle_force, 3)

    def deformation_gradients_device(self) -> torch.Tensor:
        return self._plane_device(self.F, 9).view(-1, 3, 3)

    def active_device(self) -> torch.Tensor:
        return self.input_order_device(self.active[: self.num_particles])

    def positions_numpy(self) -> np.ndarray:
        return self.positions_device().cpu().numpy()

    def velocities_numpy(self) -> np.ndarray:
        return self.velocities_device().cpu().numpy()

    def particle_forces_numpy(self) -> np.ndarray:
        return self.particle_forces_device().cpu().numpy()

    def deformation_gradients_numpy(self) -> np.ndarray:
        return self.deformation_gradients_device().cpu().numpy()

    def active_numpy(self) -> np.ndarray:
        return self.active_device().cpu().numpy()

    def masses_numpy(self) -> np.ndarray:
        return self._input_order(self.mass[: self.num_particles].cpu().numpy().copy())

    def particle_sums(self, valid_weight: torch.Tensor | None = None) -> torch.Tensor:
        """FP64 sums over the first particle_count() particles, on the device
        (10 doubles): sum m, sum m x (3), max |v|^2, particle force summed
        (3), min and sum of `valid_weight` (when given). No host copy."""
        p = self.params(self.dt_default)
        nat.call("fsi_reduce_particles", nat.ref(p), nat.ref(self._buffers), nat.ptr(valid_weight),
                 nat.ptr(self._sum_ws), self._sum_ws.numel(), nat.ptr(self._sum_out))
        return self._sum_out[:10]

    # center_of_mass / max_velocity_norm keep the reference's float32 host
    # reductions (its validation values depend on their rounding: the
    # gravity case's centre-of-mass drop is float32-summation dominated);
    # the per-step diagnostics use particle_sums().

    def center_of_mass(self) -> np.ndarray:
        # fsi/mpm3d.py:353-362
        count = self.particle_count()
        if count == 0:
            return np.zeros(3, dtype=self._np_real)
        positions = self.positions_numpy()
        masses = self.masses_numpy()
        total_mass = masses.sum()
        if total_mass <= 0.0:
            return np.zeros(3, dtype=self._np_real)
        return (positions * masses[:, None]).sum(axis=0) / total_mass

    def _com_from(self, sums: np.ndarray) -> np.ndarray:
        """Centre of mass from particle_sums() copied to the host."""
        if not sums[0] > 0.0:
            return np.zeros(3, dtype=self._np_real)
        return (sums[1:4] / sums[0]).astype(self._np_real)

    def max_velocity_norm(self) -> float:
        # fsi/mpm3d.py:364-368
        velocities = self.velocities_numpy()
        if len(velocities) == 0:
            return 0.0
        return float(np.linalg.norm(velocities, axis=1).max())
