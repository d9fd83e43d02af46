"""Learned closure on the GPU: the MFMA kernels against the CPU oracle
(oracle/mlp_oracle.py, pinned by tests/test_closure_oracle.py), the feature
kernel against the oracle's restatement, the coupler hook, and what a trained
network gains over the plain penalty force.

Tolerances: the kernels round exactly where the oracle rounds (BF16 weights,
inputs, activations and backward deltas) but accumulate in FP32 and evaluate
GELU in FP32, so a value near a BF16 rounding boundary can land one BF16 ulp
(2^-8 relative) away; 1e-2 of the largest entry covers that, while a wrong
transpose, block or bias gives O(1) errors. Features are computed in FP32:
1e-5 of each feature's largest magnitude.
"""

import dataclasses

import numpy as np
import pytest

from oracle import mlp_oracle as mo
from parity import rel_inf

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("gpu_available")]


def _data(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, mo.IN)).astype(np.float32), rng.standard_normal((n, mo.OUT)).astype(np.float32)


@pytest.mark.parametrize("n", [1, 16, 37, 1000])
def test_forward_matches_oracle(n):
    from fsi.closure import ClosureMLP

    p = mo.init_params(5)
    p[mo.O_B1:mo.O_W2] = np.linspace(-0.2, 0.2, mo.HID)
    p[mo.O_B3:mo.O_B3 + mo.OUT] = [0.1, -0.2, 0.3]
    X, _ = _data(n, 1)
    got = ClosureMLP(p).forward(X).cpu().numpy()
    ref, _ = mo.forward(p, X, bf16_round=True)
    err = rel_inf(got, ref)
    print(f"forward n={n}: rel {err:.3e}")
    assert got.shape == (n, mo.OUT) and err <= 1e-2


@pytest.mark.parametrize("n,use_idx", [(16, False), (203, False), (4096, True)])
def test_gradient_matches_oracle(n, use_idx):
    import torch

    from fsi.closure import ClosureMLP

    p = mo.init_params(6)
    p[mo.O_B2:mo.O_W3] = np.linspace(-0.1, 0.1, mo.HID)
    X, T = _data(n if not use_idx else 2 * n, 2)
    idx = None
    if use_idx:
        idx_np = np.random.default_rng(3).integers(0, 2 * n, n).astype(np.int32)
        idx = torch.from_numpy(idx_np).cuda()
        Xr, Tr = X[idx_np], T[idx_np]
    else:
        Xr, Tr = X, T
    mlp = ClosureMLP(p)
    loss = mlp.gradient(X, T, idx)
    g = mlp.grads.cpu().numpy().astype(np.float64)
    ref_loss, ref_g = mo.loss_and_grad(p, Xr, Tr, bf16_round=True)
    assert loss == pytest.approx(ref_loss / (3 * n), rel=1e-2)
    for name, lo, hi in (("W1", mo.O_W1, mo.O_B1), ("b1", mo.O_B1, mo.O_W2), ("W2", mo.O_W2, mo.O_B2),
                         ("b2", mo.O_B2, mo.O_W3), ("W3", mo.O_W3, mo.O_B3), ("b3", mo.O_B3, mo.N_PARAMS)):
        err = rel_inf(g[lo:hi], ref_g[lo:hi])
        print(f"grad {name} n={n}: rel {err:.3e}")
        assert err <= 1e-2, (name, err)
    # padding stays exactly zero
    assert np.all(g[mo.O_W1:mo.O_B1].reshape(mo.HID, mo.IN_PAD)[:, mo.IN:] == 0)
    assert np.all(g[mo.O_W3 + mo.OUT * mo.HID:mo.O_B3] == 0)


def test_adam_matches_oracle_over_steps():
    from fsi.closure import ClosureMLP

    p = mo.init_params(7)
    X, T = _data(512, 4)
    mlp = ClosureMLP(p)
    ref = p.astype(np.float64)
    m = np.zeros_like(ref)
    v = np.zeros_like(ref)
    for step in range(1, 4):
        mlp.gradient(X, T)
        g = mlp.grads.cpu().numpy().astype(np.float64)
        mlp.adam(lr=1e-3)
        ref, m, v = mo.adam(ref, g, m, v, step, lr=1e-3)
        got = mlp.parameters_numpy().astype(np.float64)
        # FP32 update of the same gradients: within a few FP32 ulps of a
        # 1e-3 step
        assert np.abs(got - ref).max() <= 1e-6, (step, np.abs(got - ref).max())
        ref = got  # continue from the device's parameters (next gradients are the device's)


def test_training_lowers_the_loss():
    from fsi.closure import train_closure

    rng = np.random.default_rng(8)
    X = rng.standard_normal((20000, mo.IN)).astype(np.float32)
    # a smooth target of a few features
    T = np.stack([np.tanh(X[:, 0] - 0.5 * X[:, 3]), X[:, 1] * X[:, 2] * 0.3, np.sin(X[:, 12])], 1).astype(np.float32)
    closure, hist = train_closure(X, T, steps=400, batch=4096, lr=3e-3, seed=0)
    print(f"loss {hist[0]:.3e} -> {hist[-1]:.3e}")
    assert hist[-1] < 0.25 * hist[0]
    # predictions through the inference kernel (normalisation, envelope)
    Xt = rng.standard_normal((500, mo.IN)).astype(np.float32) * 0.5
    Tt = np.stack([np.tanh(Xt[:, 0] - 0.5 * Xt[:, 3]), Xt[:, 1] * Xt[:, 2] * 0.3, np.sin(Xt[:, 12])], 1)
    pred = closure.predict(Xt)
    assert np.sqrt(((pred - Tt) ** 2).mean()) < 0.5 * np.sqrt((Tt ** 2).mean())
    # outside the envelope: the plain force (zero correction), counted
    before = int(closure.fallbacks.item())
    far = Xt[:4].copy()
    far[:, 5] = 1e3
    assert np.all(closure.predict(far) == 0.0)
    assert int(closure.fallbacks.item()) - before == 4
    # inside it, the correction stays within the training targets' range
    n = closure.norm
    assert np.allclose(n.out_lo, ((T - n.out_mean) / n.out_std).min(0)) and np.all(n.out_hi > n.out_lo)
    lo, hi = n.out_mean + n.out_std * n.out_lo, n.out_mean + n.out_std * n.out_hi
    near = Xt[:200] * 1.8  # inside the envelope, partly beyond the data
    p = closure.predict(near)
    assert np.all(p >= lo - 1e-5 * np.abs(lo)) and np.all(p <= hi + 1e-5 * np.abs(hi))
    tight = dataclasses.replace(n, out_lo=np.full(mo.OUT, -0.05), out_hi=np.full(mo.OUT, 0.05))
    from fsi.closure import LearnedClosure

    clipped = LearnedClosure(closure.mlp, tight).predict(Xt)
    lo, hi = n.out_mean - 0.05 * n.out_std, n.out_mean + 0.05 * n.out_std
    assert np.all(clipped >= lo - 1e-5) and np.all(clipped <= hi + 1e-5)
    assert np.isclose(clipped, lo, rtol=0, atol=1e-5).any() and np.isclose(clipped, hi, rtol=0, atol=1e-5).any()


def test_features_match_oracle():
    from fsi.closure import ClosureScene, _FeatureProbe

    scene = ClosureScene.random(11, shape=(24, 12, 12), axes=(2.5, 3.5))
    sim = scene.simulation(False)
    for _ in range(3):
        sim.step()
    u = sim.lbm.velocity_numpy()
    got = _FeatureProbe(sim.coupler).features()
    solid = sim.lbm.solid_numpy()
    phi = sim.coupler.solid_volume_fraction_numpy()
    x = sim.mpm.positions_numpy()
    v = sim.mpm.velocities_numpy()
    ref = mo.closure_features(x, v, u, solid, phi)
    for k in range(mo.IN):
        scale = max(np.abs(ref[:, k]).max(), 1e-12)
        err = np.abs(got[:, k] - ref[:, k]).max() / scale
        assert err <= 1e-5 or (k in (13, 14, 15) and err <= 1e-3), (k, err)


def test_zero_closure_keeps_the_penalty_force():
    """A network whose output layer is zero (and output mean zero) adds
    exactly zero: its correction buffer is all zeros and the run stays within
    the run-to-run noise of the plain one (float atomics in P2G make two
    plain runs differ in the last bits; the noise floor is measured)."""
    from fsi.closure import ClosureMLP, ClosureScene, LearnedClosure, Normalization

    scene = ClosureScene.random(12, shape=(24, 12, 12), axes=(2.5, 3.5))
    runs = []
    for with_closure in (False, False, True):
        sim = scene.simulation(False)
        if with_closure:
            p = mo.init_params(1)
            p[mo.O_W3:] = 0.0
            norm = Normalization(np.zeros(mo.IN), np.ones(mo.IN), np.full(mo.IN, -1e30), np.full(mo.IN, 1e30),
                                 np.zeros(mo.OUT), np.ones(mo.OUT), np.full(mo.OUT, -1e30), np.full(mo.OUT, 1e30))
            sim.coupler.set_closure(LearnedClosure(ClosureMLP(p), norm))
        for _ in range(3):
            sim.step()
        if with_closure:
            assert np.all(sim.coupler.closure_force.cpu().numpy() == 0.0)
            assert sim.coupler.closure_fallback_count() == 0
        runs.append((sim.mpm.positions_numpy(), sim.mpm.particle_forces_numpy(), sim.lbm.velocity_numpy()))
    for k, name in enumerate(("x", "pforce", "u")):
        floor = rel_inf(runs[1][k], runs[0][k])
        err = rel_inf(runs[2][k], runs[0][k])
        print(f"zero closure {name}: {err:.3e} (plain-vs-plain {floor:.3e})")
        assert err <= max(4.0 * floor, 1e-6), (name, err, floor)


def test_trained_closure_gains_over_the_plain_coarse_force():
    """Scenes of one elastic ellipsoid in inflow, each run coarse and 2x
    refined: trained on six, the closure's correction on two held-out scenes
    removes part of the coarse-vs-refined per-particle force difference."""
    from fsi.closure import ClosureScene, scene_pair_data, train_closure

    shape, axes, steps = (32, 16, 16), (2.5, 3.5), 24
    train = [scene_pair_data(ClosureScene.random(s, shape=shape, axes=axes), steps, every=2, warmup=6)
             for s in range(100, 106)]
    X = np.concatenate([d[0] for d in train])
    T = np.concatenate([d[1] for d in train])
    closure, hist = train_closure(X, T, steps=600, batch=8192, lr=2e-3, seed=1)
    gains = []
    for s in (200, 201):
        Xh, Th = scene_pair_data(ClosureScene.random(s, shape=shape, axes=axes), steps, every=2, warmup=6)
        pred = closure.predict(Xh)
        rms_plain = np.sqrt((Th.astype(np.float64) ** 2).mean())
        rms_closure = np.sqrt(((Th - pred).astype(np.float64) ** 2).mean())
        gains.append(rms_closure / rms_plain)
        print(f"held-out scene {s}: rms force difference plain {rms_plain:.3e}, with closure {rms_closure:.3e}")
    print(f"training samples {len(X)}, loss {hist[0]:.3e} -> {hist[-1]:.3e}, held-out ratios {gains}")
    assert max(gains) < 0.9


def test_gradient_is_deterministic():
    """Per-wavefront partial sums added in a fixed order: the same batch
    gives the same gradient bit for bit (no float atomics)."""
    from fsi.closure import ClosureMLP

    X, T = _data(20000, 9)
    mlp = ClosureMLP(mo.init_params(2))
    mlp.gradient(X, T)
    g1 = mlp.grads.cpu().numpy().copy()
    mlp.gradient(X, T)
    np.testing.assert_array_equal(mlp.grads.cpu().numpy(), g1)


def test_step_force_conserves_momentum_with_the_lattice():
    """The step-averaged particle forces (the closure's training targets)
    sum to minus the coupling force the same substeps spread to the fluid
    (every node of a particle's support is fluid or dropped with its weight
    renormalised, so the spread conserves momentum), and the last substep's
    force is one of the averaged terms: the two stay within the relaxation
    of one step."""
    from fsi.closure import ClosureScene

    scene = ClosureScene.random(13, shape=(24, 12, 12), axes=(2.5, 3.5))
    sim = scene.simulation(False)
    sim.coupler.record_step_force()
    for _ in range(4):
        sim.step()
    fm = sim.coupler.step_force_numpy().astype(np.float64)
    total = fm.sum(0)
    fluid = -sim.coupler.total_fluid_coupling_force().astype(np.float64)
    print(f"step-mean particle force {total}, fluid reaction {fluid}")
    assert np.abs(total - fluid).max() <= 1e-4 * np.abs(fluid).max()
    last = sim.mpm.particle_forces_numpy().astype(np.float64)
    assert np.all(np.isfinite(fm)) and np.abs(fm).max() > 0
    # gamma = 0.5 per step: within a step the slip relaxes by at most e^-0.5,
    # so the last substep's force is between 0.5x and 1.5x the step mean on
    # the body as a whole
    assert 0.5 * abs(total[0]) <= abs(last[:, 0].sum()) <= 1.5 * abs(total[0])
    sim.coupler.record_step_force(False)
    sim.step()
    with pytest.raises(RuntimeError):
        sim.coupler.step_force_numpy()


def test_half_resolution_run_tracks_the_refined_run():
    """The plain coarse run of a held-out ellipsoid scene (48x24x24) against
    its 2x-refined run: mean drag (the fluid's substep-averaged force on the
    solid) and centroid displacement within 3%, the accuracy the
    half-resolution runs are held to; the closure in the loop keeps them
    there; outside the training envelope are the first step's evaluations
    (the flow starts at the inflow speed around bodies at rest) and at most
    a few percent of the rest."""
    from fsi.closure import ClosureScene, accuracy_vs_refined, scene_pair_data, train_closure

    scene = ClosureScene.random(400, shape=(48, 24, 24), axes=(3.0, 5.0))
    data = [scene_pair_data(ClosureScene.random(s, shape=(48, 24, 24), axes=(3.0, 5.0)), 30, every=2, warmup=6)
            for s in (300, 301, 302)]
    closure, _ = train_closure(np.concatenate([d[0] for d in data]), np.concatenate([d[1] for d in data]),
                               steps=400, batch=8192, lr=2e-3, seed=1)
    acc = accuracy_vs_refined(scene, 30, closure)
    print(acc)
    for name in ("plain", "closure"):
        assert acc[name]["drag_rel_err"] <= 0.03, (name, acc[name])
        assert acc[name]["displacement_rel_err"] <= 0.03, (name, acc[name])
    per_step = acc["closure"]["evaluations"] // 30
    assert acc["closure"]["fallbacks"] <= per_step + 0.03 * (acc["closure"]["evaluations"] - per_step)


def test_step_force_of_a_non_stepping_evaluation():
    """compute_coupling_forces (one coupling pass, no step) leaves the
    accumulator at that pass's force times dt_ratio: exactly the particle
    force times 0.25, whatever it held before."""
    from fsi.closure import ClosureScene

    sim = ClosureScene.random(14, shape=(24, 12, 12), axes=(2.5, 3.5)).simulation(False)
    sim.coupler.record_step_force()
    sim.step()
    sim.coupler.compute_coupling_forces(dt_ratio=0.25)
    fm = sim.coupler.step_force_numpy()
    pf = sim.mpm.particle_forces_numpy()
    assert np.abs(pf).max() > 0
    np.testing.assert_array_equal(fm, pf * np.float32(0.25))
