"""Learned closure for the penalty coupling force (no counterpart in the
reference: it corrects the force of fsi/coupling.py:333-406).

A coarse coupled run under-resolves the flow around the solid. A small MLP
(18 -> 64 -> 64 -> 3, GELU, BF16 weights and activations, FP32
accumulation) maps per-particle features of the coarse state to a correction
of the per-particle penalty force, learned from pairs of runs of the same
scene: a coarse run and a 2x-refined one. The refined run is the same flow
under acoustic scaling (dx and dt halved): lattice velocities are equal,
the viscosity and the penalty rate in lattice units double and halve,
elastic moduli and densities are unchanged, and a force in coarse lattice
units is a quarter of the same force in fine lattice units. Each coarse
particle is the parent of the 8 fine particles that fill its volume, so the
target for a coarse particle is (sum of its children's fine forces) / 4 minus
its own coarse force, both averaged over the substeps of the same physical
interval (one coarse step, two fine steps).

Everything on the device runs in the library's HIP kernels:
``fsi_closure_features`` (features), ``fsi_mlp_train_step`` (fused forward,
loss and backward on the matrix cores), ``fsi_mlp_adam``, ``fsi_mlp_infer``
(correction, or 0 outside the training envelope). The coupler adds the
correction to the penalty force of every coupled particle
(``LBMMpmCoupler.set_closure``).
"""

from __future__ import annotations

import ctypes
import dataclasses
import functools
import math
from pathlib import Path

import numpy as np
import torch

from . import _native as nat

FEATURES = ("u_rel_x", "u_rel_y", "u_rel_z", "du_x/dx", "du_x/dy", "du_x/dz", "du_y/dx", "du_y/dy", "du_y/dz",
            "du_z/dx", "du_z/dy", "du_z/dz", "phi", "normal_x", "normal_y", "normal_z", "support", "|u_rel|")
IN, HID, OUT, N_PARAMS, N_NORM = nat.MLP_IN, nat.MLP_HIDDEN, nat.MLP_OUT, nat.MLP_PARAMS, nat.MLP_NORM
IN_PAD, OUT_PAD = 32, 16
_O_W1 = 0
_O_B1 = _O_W1 + HID * IN_PAD
_O_W2 = _O_B1 + HID
_O_B2 = _O_W2 + HID * HID
_O_W3 = _O_B2 + HID
_O_B3 = _O_W3 + OUT_PAD * HID


def init_parameters(seed: int = 0) -> np.ndarray:
    """He-uniform weights (output layer scaled by 0.1), zero biases, zero
    padding; the flat FP32 layout of the kernels."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS, np.float32)
    w1 = np.zeros((HID, IN_PAD), np.float32)
    w1[:, :IN] = rng.uniform(-1.0, 1.0, (HID, IN)) * math.sqrt(6.0 / IN)
    w3 = np.zeros((OUT_PAD, HID), np.float32)
    w3[:OUT] = rng.uniform(-1.0, 1.0, (OUT, HID)) * math.sqrt(6.0 / HID) * 0.1
    p[_O_W1:_O_B1] = w1.reshape(-1)
    p[_O_W2:_O_B2] = (rng.uniform(-1.0, 1.0, (HID, HID)) * math.sqrt(6.0 / HID)).reshape(-1)
    p[_O_W3:_O_B3] = w3.reshape(-1)
    return p


class ClosureMLP:
    """Parameters, gradients and Adam moments of the MLP on the device."""

    def __init__(self, params: np.ndarray | None = None, seed: int = 0, device: torch.device | None = None):
        self.device = device or nat.device()
        p = init_parameters(seed) if params is None else np.asarray(params, np.float32)
        if p.shape != (N_PARAMS,):
            raise ValueError(f"MLP parameters must have shape ({N_PARAMS},), got {p.shape}.")
        self.params = torch.from_numpy(p.copy()).to(self.device)
        self.grads = torch.zeros_like(self.params)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self._loss = torch.zeros(1, dtype=torch.float64, device=self.device)
        self._ws: torch.Tensor | None = None
        self.step_count = 0

    def forward(self, X: torch.Tensor) -> torch.Tensor:
        """Outputs (n x 3, normalised units) of normalised features X (n x 18)."""
        X = _rows(X, IN, self.device)
        Y = torch.empty((X.shape[0], OUT), dtype=torch.float32, device=self.device)
        nat.call("fsi_mlp_forward", X.shape[0], nat.ptr(X), nat.ptr(self.params), nat.ptr(Y))
        return Y

    def gradient(self, X: torch.Tensor, T: torch.Tensor, idx: torch.Tensor | None = None) -> float:
        """Fills self.grads with d(mean squared error)/dparams over the rows
        idx of (X, T) (all rows when None); returns the mean squared error."""
        X = _rows(X, IN, self.device)
        T = _rows(T, OUT, self.device)
        n = X.shape[0] if idx is None else int(idx.numel())
        need = ctypes.c_int64(0)
        if nat.load_library().fsi_mlp_train_workspace(n, ctypes.byref(need)) != 0:
            raise RuntimeError(nat.load_library().fsi_last_error().decode(errors="replace"))
        if self._ws is None or self._ws.numel() < need.value:
            self._ws = torch.empty(int(need.value), dtype=torch.float32, device=self.device)
        nat.fill_zero(self._loss)
        nat.call("fsi_mlp_train_step", n, nat.ptr(X), nat.ptr(T), nat.ptr(idx), nat.ptr(self.params),
                 nat.ptr(self.grads), nat.ptr(self._loss), nat.ptr(self._ws), self._ws.numel())
        return float(self._loss.item()) / (3.0 * max(n, 1))

    def adam(self, lr: float = 1.0e-3, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1.0e-8) -> None:
        self.step_count += 1
        nat.call("fsi_mlp_adam", nat.ptr(self.params), nat.ptr(self.grads), nat.ptr(self.m), nat.ptr(self.v),
                 self.step_count, float(lr), float(beta1), float(beta2), float(eps))

    def parameters_numpy(self) -> np.ndarray:
        return self.params.cpu().numpy().copy()


def _rows(t, ncols, device) -> torch.Tensor:
    t = torch.as_tensor(t, dtype=torch.float32, device=device).contiguous()
    if t.dim() != 2 or t.shape[1] != ncols:
        raise ValueError(f"expected an (n, {ncols}) array, got {tuple(t.shape)}.")
    return t


@dataclasses.dataclass
class Normalization:
    """Feature/target standardisation, the training envelope of the features
    and the range of the (normalised) training targets, which bounds what
    inference returns."""

    mean: np.ndarray
    std: np.ndarray
    lo: np.ndarray
    hi: np.ndarray
    out_mean: np.ndarray
    out_std: np.ndarray
    out_lo: np.ndarray
    out_hi: np.ndarray

    @classmethod
    def fit(cls, X: np.ndarray, T: np.ndarray, margin: float = 0.1) -> "Normalization":
        X = np.asarray(X, np.float64)
        T = np.asarray(T, np.float64)
        mean, std = X.mean(0), X.std(0)
        std = np.where(std > 1e-12, std, 1.0)
        Z = (X - mean) / std
        lo, hi = Z.min(0), Z.max(0)
        pad = margin * (hi - lo)
        omean, ostd = T.mean(0), T.std(0)
        ostd = np.where(ostd > 1e-20, ostd, 1.0)
        Y = (T - omean) / ostd
        return cls(mean, std, lo - pad, hi + pad, omean, ostd, Y.min(0), Y.max(0))

    def packed(self) -> np.ndarray:
        """The kernels' 84-float block: mean, 1/std, lo, hi, out mean, out
        std, out lo, out hi."""
        return np.concatenate([self.mean, 1.0 / self.std, self.lo, self.hi, self.out_mean,
                               self.out_std, self.out_lo, self.out_hi]).astype(np.float32)

    def inputs(self, X) -> np.ndarray:
        return ((np.asarray(X, np.float64) - self.mean) / self.std).astype(np.float32)

    def targets(self, T) -> np.ndarray:
        return ((np.asarray(T, np.float64) - self.out_mean) / self.out_std).astype(np.float32)


class LearnedClosure:
    """A trained MLP and its normalisation, evaluated per coupled step."""

    def __init__(self, mlp: ClosureMLP, norm: Normalization):
        self.mlp = mlp
        self.norm = norm
        self._norm_dev = torch.from_numpy(norm.packed()).to(mlp.device)
        self.fallbacks = torch.zeros(1, dtype=torch.int32, device=mlp.device)
        self._X: torch.Tensor | None = None

    # -------------------------------------------------------- coupler --

    def features(self, coupler) -> torch.Tensor:
        """Feature rows (capacity x 18, slot order) of the coupler's particles."""
        mpm = coupler.mpm
        cap = mpm.max_particles
        if self._X is None or self._X.shape[0] != cap:
            self._X = torch.zeros((cap, IN), dtype=torch.float32, device=self.mlp.device)
        mp = mpm.params(mpm.dt_default)
        nat.call("fsi_closure_features", nat.ref(mp), nat.ref(mpm._buffers), nat.ref(coupler._cbuf),
                 nat.ptr(self._X))
        return self._X

    def apply(self, coupler, out: torch.Tensor) -> None:
        """Writes the correction of every particle into `out` (3 x capacity,
        the coupler's real type; 0 outside the envelope, counted)."""
        X = self.features(coupler)
        n = coupler.mpm.num_particles
        if out.dtype == torch.float32:
            dst = out
        else:
            dst = torch.zeros(3 * coupler.mpm.max_particles, dtype=torch.float32, device=out.device)
        nat.call("fsi_mlp_infer", n, nat.ptr(X), nat.ptr(self.mlp.params), nat.ptr(self._norm_dev), nat.ptr(dst),
                 coupler.mpm.max_particles, nat.ptr(self.fallbacks))
        if dst is not out:
            out.copy_(dst.to(out.dtype))

    def predict(self, X_raw) -> np.ndarray:
        """Corrections (n x 3) for raw feature rows (host arrays), with the
        envelope fallback."""
        X = _rows(X_raw, IN, self.mlp.device)
        n = X.shape[0]
        out = torch.zeros(3 * max(n, 1), dtype=torch.float32, device=self.mlp.device)
        nat.call("fsi_mlp_infer", n, nat.ptr(X), nat.ptr(self.mlp.params), nat.ptr(self._norm_dev), nat.ptr(out),
                 max(n, 1), nat.ptr(self.fallbacks))
        return out.view(3, -1)[:, :n].cpu().numpy().T.copy()

    # ------------------------------------------------------ save/load --

    def save(self, path) -> Path:
        path = Path(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        n = self.norm
        np.savez(path, params=self.mlp.parameters_numpy(), mean=n.mean, std=n.std, lo=n.lo, hi=n.hi,
                 out_mean=n.out_mean, out_std=n.out_std, out_lo=n.out_lo, out_hi=n.out_hi)
        return path

    @classmethod
    def load(cls, path) -> "LearnedClosure":
        with np.load(Path(path), allow_pickle=False) as d:
            norm = Normalization(d["mean"], d["std"], d["lo"], d["hi"], d["out_mean"], d["out_std"], d["out_lo"],
                                 d["out_hi"])
            return cls(ClosureMLP(d["params"]), norm)


def train_closure(X: np.ndarray, T: np.ndarray, steps: int = 2000, batch: int = 65536, lr: float = 1.0e-3,
                  seed: int = 0, log_every: int = 0) -> tuple[LearnedClosure, list[float]]:
    """Fits the MLP to raw features X (n x 18) and targets T (n x 3) with
    Adam on random batches; returns the closure and the loss history
    (normalised mean squared error per step)."""
    norm = Normalization.fit(X, T)
    mlp = ClosureMLP(seed=seed)
    Xd = torch.from_numpy(norm.inputs(X)).to(mlp.device)
    Td = torch.from_numpy(norm.targets(T)).to(mlp.device)
    n = Xd.shape[0]
    gen = torch.Generator(device=mlp.device)
    gen.manual_seed(seed)
    b = min(batch, n)
    history = []
    for s in range(steps):
        idx = torch.randint(0, n, (b,), device=mlp.device, dtype=torch.int32, generator=gen)
        loss = mlp.gradient(Xd, Td, idx)
        mlp.adam(lr)
        history.append(loss)
        if log_every and (s % log_every == 0 or s == steps - 1):
            print(f"closure training step {s}: loss {loss:.4e}", flush=True)
    return LearnedClosure(mlp, norm), history


# ------------------------------------------------------------ scenes --


def refine_particles(parents: np.ndarray) -> np.ndarray:
    """The 8 particles of the 2x-refined run that fill each coarse particle's
    volume (coarse spacing 0.5): 2 x parent +- 0.25 in fine units; children
    of parent i at rows 8 i .. 8 i + 7."""
    off = np.array([[sx, sy, sz] for sx in (-0.25, 0.25) for sy in (-0.25, 0.25) for sz in (-0.25, 0.25)])
    p = np.asarray(parents, np.float64)
    return (2.0 * p[:, None, :] + off[None, :, :]).reshape(-1, 3).astype(np.float32)


def stable_substeps(youngs_modulus: float, poisson_ratio: float, density: float, courant: float = 0.3,
                    minimum: int = 112) -> int:
    """MPM substeps per lattice step (lattice dt = dx = 1) that keep the
    elastic P-wave Courant number c_p dt / dx <= `courant`, c_p =
    sqrt((lambda + 2 mu) / density), and at least `minimum`. The refined run
    of a pair keeps c_p in cells per lattice step (acoustic scaling), so it
    takes the same count per (fine) lattice step."""
    E, nu = float(youngs_modulus), float(poisson_ratio)
    p_modulus = E * (1.0 - nu) / ((1.0 + nu) * (1.0 - 2.0 * nu))
    return max(int(minimum), int(math.ceil(math.sqrt(p_modulus / float(density)) / courant)))


def pair_config(shape, refined: bool, n_particles: int, youngs_modulus: float, poisson_ratio: float,
                inflow: float, viscosity: float = 0.02, gamma: float = 0.5, substeps: int = 112,
                density: float = 1.2, storage: str = "fp32"):
    """SimulationConfig of a channel (velocity inflow on x-, pressure outflow
    on x+, fluid starting at the inflow velocity) for the coarse run or its
    2x refinement (acoustic scaling: viscosity x 2, penalty rate / 2, same
    moduli, density and lattice velocities; the MPM boundary layer keeps its
    physical width, 2 coarse cells = 4 fine cells, so a body reaches it at
    the same time in both runs)."""
    from .config import BoundaryConfig, CouplingConfig, LBMConfig, MPMConfig, SimulationConfig

    f = 2 if refined else 1
    nx, ny, nz = (f * int(s) for s in shape)
    lbm = LBMConfig(nx=nx, ny=ny, nz=nz, viscosity=f * viscosity, rho0=1.0, storage=storage,
                    initial_velocity=(inflow, 0.0, 0.0),
                    x_left=BoundaryConfig("velocity", velocity=(inflow, 0.0, 0.0)),
                    x_right=BoundaryConfig("pressure", rho=1.0))
    mpm = MPMConfig(nx=nx, ny=ny, nz=nz, max_particles=int(n_particles), dx=1.0, dt=1.0 / substeps,
                    density=density, youngs_modulus=youngs_modulus, poisson_ratio=poisson_ratio,
                    gravity=(0.0, 0.0, 0.0), boundary_width=2 * f)
    cpl = CouplingConfig(gamma=gamma / f, mpm_substeps_per_lbm_step=substeps)
    return SimulationConfig(num_steps=1, lbm_dt=1.0, lbm=lbm, mpm=mpm, coupling=cpl)


def pair_simulation(cfg, positions):
    """A simulation of `cfg` with channel walls and the given particles
    (volume 0.125 each, mass = density x volume)."""
    from .scenes import channel_walls
    from .simulation import FSISimulation

    sim = FSISimulation(cfg)
    sim.initialize_lbm(channel_walls((cfg.lbm.nx, cfg.lbm.ny, cfg.lbm.nz)))
    sim.initialize_mpm_from_numpy(positions, particle_volume=0.125)
    return sim


@dataclasses.dataclass(frozen=True)
class ClosureScene:
    """One seeded scene of the closure data set: a single elastic ellipsoid
    in a channel with inflow (the coarse description; `refined` doubles it)."""

    shape: tuple[int, int, int] = (48, 24, 24)
    center: tuple[float, float, float] = (16.0, 12.0, 12.0)
    semi_axes: tuple[float, float, float] = (4.0, 3.0, 3.0)
    rotation: tuple[float, ...] = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0)
    youngs_modulus: float = 1.0e3
    poisson_ratio: float = 0.3
    inflow: float = 0.03
    viscosity: float = 0.02
    gamma: float = 0.5
    substeps: int = 112
    density: float = 1.2
    seed: int = 0

    @classmethod
    def random(cls, seed: int, shape=(48, 24, 24), axes=(3.0, 5.0), modulus=(2.0e2, 2.0e3),
               poisson=(0.2, 0.4), inflow=(0.01, 0.05), substeps: int | str = 112) -> "ClosureScene":
        """Semi-axes U[axes], random orientation, E log-uniform in `modulus`,
        nu U[poisson], inflow U[inflow]; centred a third along the channel.
        substeps="stable": the count ``stable_substeps`` gives the scene's
        material."""
        from .scenes import random_rotation

        rng = np.random.default_rng(seed)
        semi = tuple(float(a) for a in rng.uniform(axes[0], axes[1], 3))
        R = random_rotation(rng)
        nx, ny, nz = shape
        E = float(math.exp(rng.uniform(math.log(modulus[0]), math.log(modulus[1]))))
        nu = float(rng.uniform(*poisson))
        inflow_x = float(rng.uniform(*inflow))
        n_sub = stable_substeps(E, nu, cls.density) if substeps == "stable" else int(substeps)
        return cls(shape=tuple(shape), center=(nx / 3.0, ny / 2.0, nz / 2.0), semi_axes=semi,
                   rotation=tuple(float(v) for v in R.reshape(-1)), youngs_modulus=E, poisson_ratio=nu,
                   inflow=inflow_x, substeps=n_sub, seed=seed)

    def particles(self) -> tuple[np.ndarray, np.ndarray]:
        """Coarse particles (spacing 0.5, jittered) inside the ellipsoid and
        their children for the refined run (``refine_particles``)."""
        rng = np.random.default_rng(self.seed + 7919)
        c = np.asarray(self.center)
        semi = np.asarray(self.semi_axes)
        R = np.asarray(self.rotation).reshape(3, 3)
        ext = np.abs(R) @ semi
        axes = [np.arange(np.floor(c[a] - ext[a]) + 0.25, c[a] + ext[a], 0.5) for a in range(3)]
        grid = np.stack([m.ravel() for m in np.meshgrid(*axes, indexing="ij")], axis=1)
        local = (grid - c) @ R
        parents = grid[((local / semi) ** 2).sum(1) <= 1.0]
        parents = (parents + rng.uniform(-0.125, 0.125, parents.shape)).astype(np.float32)
        return parents, refine_particles(parents)

    def config(self, refined: bool = False, storage: str = "fp32"):
        """SimulationConfig of the coarse run or of the 2x-refined run."""
        npart = len(self.particles()[1 if refined else 0])
        return pair_config(self.shape, refined, npart, self.youngs_modulus, self.poisson_ratio, self.inflow,
                           self.viscosity, self.gamma, self.substeps, self.density, storage)

    def simulation(self, refined: bool = False, storage: str = "fp32"):
        parents, children = self.particles()
        return pair_simulation(self.config(refined, storage), children if refined else parents)

    def bodies(self) -> list[tuple[int, int]]:
        return [(0, len(self.particles()[0]))]


class _FeatureProbe:
    """Feature rows of a coupler's particles in input order (host)."""

    def __init__(self, coupler):
        self.coupler = coupler
        self.X = torch.zeros((coupler.mpm.max_particles, IN), dtype=torch.float32, device=coupler.mpm.device)

    def features(self) -> np.ndarray:
        self.coupler.lbm._ensure_macro()
        mpm = self.coupler.mpm
        mp = mpm.params(mpm.dt_default)
        nat.call("fsi_closure_features", nat.ref(mp), nat.ref(mpm._buffers), nat.ref(self.coupler._cbuf),
                 nat.ptr(self.X))
        return mpm._input_order(self.X[: mpm.num_particles].cpu().numpy().copy())


def scene_pair_data(scene, steps: int, every: int = 4, warmup: int = 8) -> tuple[np.ndarray, np.ndarray]:
    """Runs the coarse and refined simulations of a scene side by side and
    returns (features, targets) of the coarse particles for coarse steps
    warmup, warmup + every, ... < steps: the features of the state at the
    start of the step and the difference of the step-averaged coupling
    forces over that step (the refined run's two steps: the mean of their
    step averages, children summed, in coarse units). The correction is
    applied on every substep of a coarse step, so its target is the
    difference of the forces averaged over the step, not of the last
    substep's forces (a point sample of a force that relaxes within a step
    twice as long as a fine one)."""
    coarse = scene.simulation(False)
    fine = scene.simulation(True)
    coarse.coupler.record_step_force()
    fine.coupler.record_step_force()
    probe = _FeatureProbe(coarse.coupler)
    Xs, Ts = [], []
    for k in range(steps):
        take = k >= warmup and (k - warmup) % every == 0
        if take:
            X = probe.features()
            active = coarse.mpm.active_numpy() != 0
        coarse.step()
        ff = 0.0
        for _ in range(2):
            fine.step()
            if take:
                ff = ff + fine.coupler.step_force_numpy().astype(np.float64)
        if take:
            fc = coarse.coupler.step_force_numpy().astype(np.float64)
            ff = ff.reshape(-1, 8, 3).sum(1) / (2.0 * 4.0)
            Xs.append(X[active])
            Ts.append((ff - fc)[active])
    return np.concatenate(Xs).astype(np.float32), np.concatenate(Ts).astype(np.float32)


@dataclasses.dataclass(frozen=True)
class PackScene:
    """Several elastic spheres in a channel with inflow (the coarse
    description; `refined` doubles it), the multi-body counterpart of
    ``ClosureScene``."""

    shape: tuple[int, int, int] = (48, 24, 24)
    count: int = 4
    radii: tuple[float, float] = (2.5, 5.0)
    youngs_modulus: float = 1.0e3
    poisson_ratio: float = 0.3
    inflow: float = 0.04
    viscosity: float = 0.02
    gamma: float = 0.5
    substeps: int = 112
    density: float = 1.2
    margin: float = 3.0
    gap: float = 1.5
    upstream: float = 1.0
    seed: int = 0
    # per-sphere Young's modulus, log-uniform in this range (a per-particle
    # material plane; the config's global modulus is then unused), or None
    body_modulus: tuple[float, float] | None = None

    @classmethod
    def random(cls, seed: int, shape=(48, 24, 24), count=4, radii=(2.5, 5.0), modulus=(2.0e2, 2.0e3),
               poisson=(0.2, 0.4), inflow=(0.01, 0.05), upstream=0.7) -> "PackScene":
        """E log-uniform in `modulus`, nu U[poisson], inflow U[inflow]; the
        spheres are packed into the first `upstream` fraction of the
        channel, so that none drifts into the outflow's boundary layer."""
        rng = np.random.default_rng(seed + 104729)
        return cls(shape=tuple(shape), count=count, radii=tuple(radii), upstream=float(upstream),
                   youngs_modulus=float(math.exp(rng.uniform(math.log(modulus[0]), math.log(modulus[1])))),
                   poisson_ratio=float(rng.uniform(*poisson)), inflow=float(rng.uniform(*inflow)), seed=seed)

    def _pack(self):
        return _packed_spheres(self)

    def particles(self) -> tuple[np.ndarray, np.ndarray]:
        parents, _ = self._pack()
        return parents, refine_particles(parents)

    def bodies(self) -> list[tuple[int, int]]:
        """Coarse particle range [start, stop) of each sphere."""
        counts = self._pack()[1]["counts"]
        ends = np.cumsum(counts)
        return [(int(e - c), int(e)) for c, e in zip(counts, ends)]

    def body_moduli(self) -> np.ndarray | None:
        """Young's modulus of each sphere (``body_modulus``: log-uniform,
        numpy default_rng(seed + 7)), or None."""
        if self.body_modulus is None:
            return None
        from .scenes import body_moduli

        return body_moduli(self._pack()[1]["counts"], self.body_modulus[0], self.body_modulus[1],
                           seed=self.seed + 7)[1]

    def config(self, refined: bool = False, storage: str = "fp32"):
        npart = len(self.particles()[1 if refined else 0])
        E = self.youngs_modulus if self.body_modulus is None else float(self.body_moduli().max())
        return pair_config(self.shape, refined, npart, E, self.poisson_ratio, self.inflow,
                           self.viscosity, self.gamma, self.substeps, self.density, storage)

    def simulation(self, refined: bool = False, storage: str = "fp32"):
        parents, children = self.particles()
        sim = pair_simulation(self.config(refined, storage), children if refined else parents)
        if self.body_modulus is not None:
            per_particle = np.repeat(self.body_moduli(), np.asarray(self._pack()[1]["counts"], np.int64))
            sim.mpm.set_particle_material(np.repeat(per_particle, 8) if refined else per_particle)
        return sim


@functools.lru_cache(maxsize=8)
def _packed_spheres(scene: PackScene):
    from .scenes import sphere_pack

    box = (int(round(scene.shape[0] * scene.upstream)),) + tuple(scene.shape[1:])
    return sphere_pack(box, scene.count, scene.radii[0], scene.radii[1], spacing=0.5, seed=scene.seed,
                       margin=scene.margin, gap=scene.gap)


def run_observables(sim, steps: int, bodies, refined: bool = False, force_every: int = 0,
                    progress: str | None = None) -> dict:
    """Runs `steps` coarse steps (twice as many fine steps when `refined`)
    and returns, in coarse lattice units, the drag per coarse step, each
    body's centroid displacement and, every `force_every` coarse steps, the
    step-averaged coupling force of every coarse particle (for the refined
    run: its 8 children summed, mean of the two fine steps).

    The drag is the x-force of the fluid on the solids: the step's coupling
    force summed over the lattice (an FP64 device sum), sign flipped. It is
    the substep-averaged force (the coupling buffer accumulates every
    substep's spread force times its time fraction), so it is the momentum
    the fluid gives the solids per unit time. The end-of-step particle force
    is a point sample of a force that relaxes within the step, and a fine
    step is half as long, so that sample does not compare across
    resolutions. A fine force is a quarter of the same force in coarse
    units; the fine drag of a coarse step is the mean of its two fine steps.
    `bodies` are coarse particle ranges (children of coarse particle i are
    fine particles 8 i .. 8 i + 7). `progress`: a label; a line on stderr
    every 100 coarse steps."""
    import sys
    import time

    f = 2 if refined else 1
    t0 = time.perf_counter()
    rows = [(f ** 3 * a, f ** 3 * b) for a, b in bodies]
    if force_every:
        sim.coupler.record_step_force()

    def centroids():
        x = sim.mpm.positions_numpy().astype(np.float64) / f
        return np.stack([x[a:b].mean(0) for a, b in rows])

    c0 = centroids()
    drag = np.zeros(steps)
    forces = []
    for k in range(steps):
        take = force_every and (k + 1) % force_every == 0
        pf = 0.0
        for _ in range(f):
            sim.step()
            drag[k] -= float(sim.lbm.lattice_sums(sim.coupler.cforce)[2])
            if take:
                pf = pf + sim.coupler.step_force_numpy().astype(np.float64)
        drag[k] /= f * f * f
        if progress and (k + 1) % 100 == 0:
            print(f"[{progress}] {k + 1}/{steps} coarse steps, {time.perf_counter() - t0:.0f} s", file=sys.stderr,
                  flush=True)
        if take:
            forces.append(pf.reshape(-1, f ** 3, 3).sum(1) / (f * f * f) if refined else pf)
    return {"drag_x": drag, "displacement": centroids() - c0, "particle_forces": forces,
            "fallbacks": sim.coupler.closure_fallback_count() if sim.coupler.closure is not None else 0}


def observable_errors(run: dict, fine: dict) -> dict:
    """Relative errors of a coarse run against the refined run: mean drag
    over the whole run and over its second half, the displacement error
    |d - d_fine| / |d_fine| (vectors) averaged over the bodies and, when
    both runs sampled them, the RMS per-particle step-averaged force error
    relative to the RMS refined force (the quantity the closure corrects,
    here with the closure in the loop)."""
    d, df = run["drag_x"], fine["drag_x"]
    h = len(d) // 2
    disp = np.linalg.norm(run["displacement"] - fine["displacement"], axis=1) / np.maximum(
        np.linalg.norm(fine["displacement"], axis=1), 1e-30)
    out = {"drag_rel_err": round(float(abs(d.mean() / df.mean() - 1.0)), 4),
           "drag_rel_err_second_half": round(float(abs(d[h:].mean() / df[h:].mean() - 1.0)), 4),
           "displacement_rel_err": round(float(disp.mean()), 4),
           "displacement_rel_err_max": round(float(disp.max()), 4)}
    pa, pb = run.get("particle_forces") or [], fine.get("particle_forces") or []
    if pa and len(pa) == len(pb):
        num = sum(float(((a - b) ** 2).sum()) for a, b in zip(pa, pb))
        den = sum(float((b ** 2).sum()) for b in pb)
        out["particle_force_rel_rms"] = round(float(np.sqrt(num / max(den, 1e-300))), 4)
    return out


def accuracy_vs_refined(scene, steps: int, closure: LearnedClosure | None = None, storage: str = "fp32",
                        force_every: int = 5, progress: str | None = None, plain_storages=()) -> dict:
    """Drag, displacement and per-particle force errors of the coarse run,
    plain and with the closure (coarse populations stored as `storage`),
    against the refined FP32 run of the same scene (forces sampled every
    `force_every` coarse steps); the closure run's fallback count (particle
    evaluations outside the training envelope) and evaluations.
    `plain_storages`: further plain coarse runs with other population
    storages ("plain_<storage>")."""
    bodies = scene.bodies()
    fine = run_observables(scene.simulation(True), steps, bodies, refined=True, force_every=force_every,
                           progress=progress and progress + " refined")
    out = {"fine_mean_drag_x": float(fine["drag_x"].mean())}
    runs = [("plain", None, storage)] + ([("closure", closure, storage)] if closure is not None else [])
    runs += [(f"plain_{st}", None, st) for st in plain_storages]
    for name, cl, st in runs:
        sim = scene.simulation(False, st)
        if cl is not None:
            sim.coupler.set_closure(cl)
        r = run_observables(sim, steps, bodies, force_every=force_every,
                            progress=progress and f"{progress} coarse {name}")
        out[name] = observable_errors(r, fine)
        if cl is not None:
            out[name]["fallbacks"] = r["fallbacks"]
            out[name]["evaluations"] = int(steps * sim.mpm.num_particles)
    return out
