"""Explicit partitioned penalty coupling on the GPU (reference: fsi/coupling.py).

Same constructor, methods, accessors and exceptions as the reference's
``LBMMpmCoupler`` (fsi/coupling.py:11-496). The coupling kernels are fused
into the MPM passes: velocity interpolation, penalty force, reaction and
occupancy spreading run inside P2G, contact damping inside G2P, and the
occupancy clamp, lattice force reset, coupling-force application and
immersed-boundary drag are one cell pass before the lattice step.

Per coupled step (fsi/coupling.py:135-173):
  clear coupling fields
  S x [clear grid; P2G + coupling; grid update + G2P (+ contact)]
  force assembly (clamp phi, reset force, add coupling force, IB drag)
  lattice step
"""

from __future__ import annotations

import numpy as np
import torch

from . import _native as nat
from .config import CouplingConfig
from .lbm3d import LBMSolver3D
from .mpm3d import MPMSolver3D

# counter slots of the device counter array
_UNSUPPORTED, _PARTIAL, _CLIPPED, _CONTACT, _DAMPED, _IB_ACTIVE, _IB_CLIPPED = range(7)


class LBMMpmCoupler:
    """Explicit partitioned penalty coupling between the GPU LBM and MPM solvers."""

    def __init__(self, config: CouplingConfig, lbm: LBMSolver3D, mpm: MPMSolver3D):
        self.config = config
        self.config.validate()
        self.lbm = lbm
        self.mpm = mpm
        self._validate_solver_contract()
        self.enabled = bool(config.enabled)
        self.build_solid_volume_fraction = bool(config.build_solid_volume_fraction)
        self.immersed_boundary_enabled = bool(config.immersed_boundary_enabled)
        self.contact_enabled = bool(config.contact_enabled)
        self.substeps = int(config.mpm_substeps_per_lbm_step)

        self.nx = int(lbm.nx)
        self.ny = int(lbm.ny)
        self.nz = int(lbm.nz)
        # the lattice cell the coupling works on: the MPM cell (the
        # reference), or mpm_refinement MPM cells per axis
        self.refine = int(config.mpm_refinement)
        self.inv_dx = 1.0 / (self.refine * float(mpm.dx))
        self.cell_volume = (self.refine * float(mpm.dx)) ** 3

        if lbm.real_dtype != mpm.real_dtype:
            raise ValueError("LBM and MPM solvers must use the same real type for coupling.")
        n = self.nx * self.ny * self.nz
        dev, rd = lbm.device, lbm.real_dtype
        self._np_real = np.float64 if rd == torch.float64 else np.float32
        # (force density xyz, phi) per cell
        self.cforce = torch.zeros(4 * n, dtype=rd, device=dev)
        self.ib_force = torch.zeros(3 * n, dtype=rd, device=dev)
        self.ib_total = torch.zeros(3, dtype=rd, device=dev)
        cap = mpm.max_particles
        self.valid_weight = torch.zeros(cap, dtype=rd, device=dev)
        self.coupling_mask = torch.zeros(cap, dtype=torch.int32, device=dev)
        self.contact_mask = torch.zeros(cap, dtype=torch.int32, device=dev)
        self.counters = torch.zeros(8, dtype=torch.int32, device=dev)
        self.coupled_step = 0
        # the particles are re-sorted by cell key every `sort_interval`
        # coupled steps (0: never), keeping each P2G workgroup's particles in
        # a compact box; results change only by float rounding of the sums
        self.sort_interval = 1
        self._steps_since_sort = 0
        # per-substep counter rows of the fused substep launches
        self._counter_rows = torch.zeros(8 * max(self.substeps, 1), dtype=torch.int32, device=dev)
        self._cbuf = nat.CouplingBuffers(
            nat.ptr(lbm.solid), nat.ptr(lbm.vel), nat.ptr(self.cforce), nat.ptr(self.valid_weight),
            nat.ptr(self.coupling_mask), nat.ptr(self.contact_mask), nat.ptr(self.counters), None,
            nat.ptr(self._counter_rows), max(self.substeps, 1),
        )
        # optional learned closure (fsi.closure): a per-particle correction to
        # the penalty force, evaluated once per coupled step
        self.closure = None
        self.closure_force: torch.Tensor | None = None
        # optional step-averaged coupling force per particle slot
        self.force_mean: torch.Tensor | None = None

    def set_closure(self, closure) -> None:
        """Attach a ``fsi.closure.LearnedClosure`` (None detaches it). Each
        coupled step then evaluates it on the state at the start of the step
        and adds its correction to every coupled particle's penalty force
        (and so to the fluid's reaction); particles outside the training
        envelope keep the plain force. Not part of the reference."""
        if closure is not None and self.refine > 1:
            raise NotImplementedError("the learned closure needs co-located grids (mpm_refinement 1).")
        self.closure = closure
        if closure is None:
            self.closure_force = None
            self._cbuf.closure_force = None
            return
        self._fallback_base = int(closure.fallbacks.item())
        cap = self.mpm.max_particles
        self.closure_force = torch.zeros(3 * cap, dtype=self.lbm.real_dtype, device=self.lbm.device)
        self._cbuf.closure_force = self.closure_force.data_ptr()

    def record_step_force(self, enabled: bool = True) -> None:
        """Keep the step-averaged coupling force of every particle (the mean
        over the step's substeps, which the reference does not store: its
        particle force is the last substep's, fsi/coupling.py:389); read it
        with ``step_force_numpy``. Not part of the reference."""
        if enabled:
            self.force_mean = torch.zeros(3 * self.mpm.max_particles, dtype=self.lbm.real_dtype,
                                          device=self.lbm.device)
            self._cbuf.force_mean = self.force_mean.data_ptr()
        else:
            self.force_mean = None
            self._cbuf.force_mean = None

    def step_force_numpy(self) -> np.ndarray:
        """(n, 3) step-averaged coupling force of the last step (after
        ``compute_coupling_forces``: its force times dt_ratio), in input
        order (``record_step_force`` must be on)."""
        if self.force_mean is None:
            raise RuntimeError("record_step_force() is off")
        n, cap = self.mpm.num_particles, self.mpm.max_particles
        f = self.force_mean.view(3, cap)[:, :n].cpu().numpy().T.copy()
        return self.mpm._input_order(f)

    def closure_fallback_count(self) -> int:
        """Particle evaluations that left the closure's training envelope (and
        kept the plain force) since the closure was attached."""
        return 0 if self.closure is None else int(self.closure.fallbacks.item()) - self._fallback_base

    def _validate_solver_contract(self) -> None:
        # fsi/coupling.py:42-56
        if self.config.kernel != "quadratic":
            raise ValueError(f"Unsupported coupling kernel: {self.config.kernel}.")
        r = int(self.config.mpm_refinement)
        lbm_dims = (self.lbm.nx, self.lbm.ny, self.lbm.nz)
        mpm_dims = (self.mpm.nx, self.mpm.ny, self.mpm.nz)
        if r == 1 and lbm_dims != mpm_dims:
            raise ValueError(
                "LBM and MPM grids must match for Step 5 coupling. "
                f"Got LBM={lbm_dims}, "
                f"MPM={mpm_dims}."
            )
        if r > 1 and mpm_dims != tuple(r * n for n in lbm_dims):
            raise ValueError(f"With mpm_refinement={r} the MPM grid must be {r} x the LBM grid. "
                             f"Got LBM={lbm_dims}, MPM={mpm_dims}.")
        if self.mpm.dx <= 0.0:
            raise ValueError("MPM dx must be positive.")

    def _validate_initialized(self) -> None:
        if not getattr(self.lbm, "_initialized", False):
            raise RuntimeError("LBMSolver3D.initialize must be called before coupling.")
        if not getattr(self.mpm, "_initialized", False):
            raise RuntimeError("MPMSolver3D particle initialization must be called before coupling.")

    # ------------------------------------------------------- parameters --

    def effective_gamma(self) -> float:
        # fsi/coupling.py:247-253 (gamma is a float32 field in the reference)
        gamma = float(np.float32(self.config.gamma))
        ramp_steps = int(self.config.gamma_ramp_steps)
        if ramp_steps <= 0:
            return gamma
        return gamma * min(1.0, float(self.coupled_step + 1) / float(ramp_steps))

    def _kernel_gamma(self) -> float:
        # fsi/coupling.py:321-330: the ramp evaluated in the kernel's float type
        rd = self._np_real
        gamma = rd(self.config.gamma)
        ramp_steps = int(self.config.gamma_ramp_steps)
        if ramp_steps > 0:
            ramp = (rd(self.coupled_step) + rd(1.0)) / rd(ramp_steps)
            gamma = gamma * min(rd(1.0), ramp)
        return float(gamma)

    def _cparams(self, dt_ratio: float, build: bool, enabled: bool) -> nat.CouplingParams:
        c = self.config
        p = nat.CouplingParams()
        p.enabled = 1 if enabled else 0
        p.build_fraction = 1 if build else 0
        p.gamma_eff = self._kernel_gamma()
        p.dt_ratio = float(self._np_real(dt_ratio))
        p.cell_volume = float(self._np_real(self.cell_volume))
        p.refine = self.refine
        p.min_valid_weight = float(c.min_valid_weight)
        p.use_relative_velocity_limit = 0 if c.relative_velocity_limit is None else 1
        p.relative_velocity_limit = 0.0 if c.relative_velocity_limit is None else float(c.relative_velocity_limit)
        p.use_force_limit = 0 if c.force_limit is None else 1
        p.force_limit = 0.0 if c.force_limit is None else float(c.force_limit)
        p.contact_enabled = 1 if self.contact_enabled else 0
        p.contact_threshold = float(c.contact_fraction_threshold)
        p.contact_damping = float(c.contact_velocity_damping)
        return p

    def _assembly(self, coupling: bool, clamp: bool) -> nat.ForceAssembly:
        c = self.config
        a = nat.ForceAssembly()
        a.coupling = 1 if coupling else 0
        a.clamp_fraction = 1 if clamp else 0
        a.ib_enabled = 1 if self.immersed_boundary_enabled else 0
        a.ib_drag = float(c.immersed_boundary_drag)
        a.ib_threshold = float(c.immersed_boundary_fraction_threshold)
        a.ib_max_force = 0.0 if c.immersed_boundary_max_force is None else float(c.immersed_boundary_max_force)
        a.ib_use_max = 0 if c.immersed_boundary_max_force is None else 1
        a.base_force = nat.c_double3(*[float(x) for x in self.lbm.config.force])
        return a

    # ------------------------------------------------------------ steps --

    def step(self, lbm_dt: float) -> None:
        """Run one explicit coupled LBM step with the configured MPM substeps."""
        lbm_dt_value = float(lbm_dt)
        if lbm_dt_value <= 0.0:
            raise ValueError("lbm_dt must be positive.")
        self._validate_initialized()

        self.lbm._ensure_macro()  # the substeps interpolate the lattice velocity of this state
        if self.sort_interval > 0 and self._steps_since_sort % self.sort_interval == 0:
            self.mpm.sort_particles()
        self._steps_since_sort += 1
        if self.closure is not None and self.enabled:
            # features read the occupancy of the last step: before the clear
            self.closure.apply(self, self.closure_force)
        if self.force_mean is not None:
            nat.fill_zero(self.force_mean)
        self.clear_coupling_fields()
        mpm_dt = lbm_dt_value / float(self.substeps)
        dt_ratio = mpm_dt / lbm_dt_value
        mp = self.mpm.params(self.mpm._dt_value(mpm_dt))
        mb = self.mpm._buffers
        cp = self._cparams(dt_ratio, self.build_solid_volume_fraction, self.enabled)
        # S x (clear grid, P2G + coupling, grid update + G2P + contact) in one call
        nat.call("fsi_mpm_substeps", nat.ref(mp), nat.ref(mb), nat.ref(cp), nat.ref(self._cbuf),
                 self.substeps, 1 if self.build_solid_volume_fraction else 0, 1)

        # clamp phi, reset the lattice force, add the coupling force and IB
        # drag, then the lattice step's collision with that force: one pass
        a = self._assembly(self.enabled, self.enabled and self.build_solid_volume_fraction)
        if self.enabled or self.immersed_boundary_enabled:
            self.lbm.collide_assembled(a, self.cforce, self.ib_force, self.ib_total, self.counters)
        else:
            # nothing couples: the force is the uniform base force
            self.lbm._before_force_update()
            nat.call("fsi_force_assemble", nat.ref(self.lbm._params), nat.ref(self.lbm._buffers),
                     nat.ref(a), nat.ptr(self.cforce), nat.ptr(self.ib_force), nat.ptr(self.ib_total),
                     nat.ptr(self.counters), 1)
            self.lbm._force_uniform = self.lbm._rounded(self.lbm.config.force)
            self.lbm.step()
        if self.enabled:
            self.coupled_step += 1

    def compute_coupling_forces(self, dt_ratio: float = 1.0) -> None:
        """Compute particle force and fluid reaction force without stepping (coupling.py:175-196)."""
        dt_ratio_value = float(dt_ratio)
        if dt_ratio_value < 0.0:
            raise ValueError("dt_ratio must be non-negative.")
        self._validate_initialized()
        self.lbm._ensure_macro()
        self.clear_coupling_fields()
        self.mpm.clear_particle_force()
        if self.force_mean is not None:
            nat.fill_zero(self.force_mean)
        if not self.enabled:
            return
        mp = self.mpm.params(self.mpm.dt_default)
        mb = self.mpm._buffers
        cp = self._cparams(dt_ratio_value, self.build_solid_volume_fraction, True)
        nat.call("fsi_mpm_p2g", nat.ref(mp), nat.ref(mb), nat.ref(cp), nat.ref(self._cbuf), 1, 0)
        a = self._assembly(False, self.build_solid_volume_fraction)
        nat.call("fsi_force_assemble", nat.ref(self.lbm._params), nat.ref(self.lbm._buffers),
                 nat.ref(a), nat.ptr(self.cforce), nat.ptr(self.ib_force), nat.ptr(self.ib_total),
                 nat.ptr(self.counters), 0)
        if self.contact_enabled:
            nat.call("fsi_contact", nat.ref(mp), nat.ref(mb), nat.ref(cp), nat.ref(self._cbuf))

    def clear_coupling_fields(self) -> None:
        # fsi/coupling.py:281-289 (+ 291-301)
        nat.fill_zero(self.cforce)
        nat.fill_zero(self.ib_force)
        nat.fill_zero(self.ib_total)
        self.clear_particle_diagnostics()

    def clear_particle_diagnostics(self) -> None:
        # fsi/coupling.py:291-301 (entries beyond the particle count are never read)
        nat.fill_zero(self.counters)
        nat.fill_zero(self.valid_weight)
        nat.fill_zero(self.coupling_mask)
        nat.fill_zero(self.contact_mask)

    # -------------------------------------------------------- accessors --

    # Device accessors: the *_numpy arrays as fresh device tensors (same
    # shapes, dtypes and particle order); the *_numpy accessors copy them.

    def _cells_device(self, t: torch.Tensor, ncomp: int) -> torch.Tensor:
        return t.view(self.nx, self.ny, self.nz, ncomp)

    def coupling_force_device(self) -> torch.Tensor:
        return self._cells_device(self.cforce, 4)[..., :3].contiguous()

    def solid_volume_fraction_device(self) -> torch.Tensor:
        return self._cells_device(self.cforce, 4)[..., 3].contiguous()

    def immersed_boundary_force_device(self) -> torch.Tensor:
        out = torch.empty_like(self.ib_force)
        nat.call("fsi_soa_to_aos", self.lbm.real_code, self.nx * self.ny * self.nz, 3,
                 nat.ptr(self.ib_force), nat.ptr(out))
        return self._cells_device(out, 3)

    def _per_particle_device(self, t: torch.Tensor) -> torch.Tensor:
        # slot order -> input order (the particles may have been sorted)
        return self.mpm.input_order_device(t[: self.mpm.particle_count()])

    def particle_valid_weights_device(self) -> torch.Tensor:
        return self._per_particle_device(self.valid_weight)

    def particle_coupling_mask_device(self) -> torch.Tensor:
        return self._per_particle_device(self.coupling_mask)

    def particle_contact_mask_device(self) -> torch.Tensor:
        return self._per_particle_device(self.contact_mask)

    def coupling_force_numpy(self) -> np.ndarray:
        return self.coupling_force_device().cpu().numpy()

    def solid_volume_fraction_numpy(self) -> np.ndarray:
        return self.solid_volume_fraction_device().cpu().numpy()

    def immersed_boundary_force_numpy(self) -> np.ndarray:
        return self.immersed_boundary_force_device().cpu().numpy()

    def particle_valid_weights_numpy(self) -> np.ndarray:
        return self.particle_valid_weights_device().cpu().numpy()

    def particle_coupling_mask_numpy(self) -> np.ndarray:
        return self.particle_coupling_mask_device().cpu().numpy()

    def particle_contact_mask_numpy(self) -> np.ndarray:
        return self.particle_contact_mask_device().cpu().numpy()

    # The accessors keep the reference's float32 host reductions;
    # diagnostic_vector() gathers the FP64 device reductions the per-step
    # diagnostics use (one small host copy instead of whole fields).

    def _real(self):
        return np.float64 if self.lbm.real_dtype == torch.float64 else np.float32

    def diagnostic_vector(self) -> np.ndarray:
        """One host copy of every reduced quantity, float64:
        [0:5] lattice sums (fluid mass, max |u|^2, fluid coupling force xyz),
        [5:15] particle sums (sum m, sum m x, max |v|^2, particle force xyz,
        min and sum of valid weights), [15:23] counters, [23:26] IB total."""
        lat = self.lbm.lattice_sums(self.cforce)
        part = self.mpm.particle_sums(self.valid_weight)
        return torch.cat([lat, part, self.counters.to(torch.float64),
                          self.ib_total.to(torch.float64)]).cpu().numpy()

    def total_particle_coupling_force(self) -> np.ndarray:
        # fsi/coupling.py:238-239
        return self.mpm.particle_forces_numpy().sum(axis=0)

    def total_fluid_coupling_force(self) -> np.ndarray:
        # fsi/coupling.py:241-244
        force_np = self.coupling_force_numpy()
        solid_np = self.lbm.solid_numpy()
        return force_np[solid_np == 0].sum(axis=0)

    def _counters(self) -> np.ndarray:
        return self.counters.cpu().numpy()

    @staticmethod
    def _ib_dict(ctr, ib_total_force) -> dict[str, int | float]:
        ib_total_force = np.asarray(ib_total_force, dtype=np.float64)
        return {
            "ib_active_cell_count": int(ctr[_IB_ACTIVE]),
            "ib_clipped_cell_count": int(ctr[_IB_CLIPPED]),
            "ib_total_force_x": float(ib_total_force[0]),
            "ib_total_force_y": float(ib_total_force[1]),
            "ib_total_force_z": float(ib_total_force[2]),
            "ib_total_force_norm": float(np.linalg.norm(ib_total_force)),
            "contact_candidate_count": int(ctr[_CONTACT]),
            "contact_damped_particle_count": int(ctr[_DAMPED]),
        }

    def immersed_boundary_diagnostics(self) -> dict[str, int | float]:
        # fsi/coupling.py:246-257
        return self._ib_dict(self._counters(), self.ib_total.cpu().numpy())

    def coupling_diagnostics(self, vector: np.ndarray | None = None) -> dict[str, int | float]:
        # fsi/coupling.py:263-276. Without `vector` the weight statistics are
        # the reference's float32 host ones; with a diagnostic_vector()
        # already on the host they come from its FP64 device sums.
        count = self.mpm.particle_count()
        if vector is None:
            weights = self.particle_valid_weights_numpy()
            ctr = self._counters()
            ib = self.ib_total.cpu().numpy()
            min_weight = float(weights.min()) if count else 0.0
            mean_weight = float(weights.mean()) if count else 0.0
        else:
            ctr = vector[15:23].astype(np.int64)
            ib = vector[23:26]
            min_weight = float(self._real()(vector[13])) if count else 0.0
            mean_weight = float(self._real()(vector[14] / count)) if count else 0.0
        return {
            "unsupported_particle_count": int(ctr[_UNSUPPORTED]),
            "partial_support_particle_count": int(ctr[_PARTIAL]),
            "clipped_particle_count": int(ctr[_CLIPPED]),
            "min_particle_valid_weight": min_weight,
            "mean_particle_valid_weight": mean_weight,
            "effective_gamma": self.effective_gamma(),
            **self._ib_dict(ctr, ib),
        }
