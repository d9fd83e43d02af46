"""The learned closure's two scene configurations at their stated sizes, on
the GPU (short runs; the long ones are bench.py --closure-config).

* Training pairs: a single elastic ellipsoid (semi-axes U[4, 10] coarse
  cells, E log-uniform in [2e2, 2e4], nu U[0.2, 0.4], inflow U[0.01, 0.05])
  in a 96x48x48 coarse channel against its 2x refinement, 192x96x96 with 8
  particles per coarse particle. Checked: the sizes, the feature and target
  rows of the sampled steps, and the force balance the targets rest on.
* Large scene: 64 elastic spheres (radius 5-14 refined cells, E log-uniform
  in [5e2, 5e3] per sphere) in a 384x192x192 FP32 channel, and its
  192x96x96 coarse description with FP16 populations. Checked: the sizes,
  the per-sphere material, the momentum balance of the penalty coupling
  (fsi/coupling.py:332-406), and that both runs stay finite and bounded.
"""

import numpy as np
import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("gpu_available")]


def test_pair_data_at_the_configuration_size():
    import bench
    from fsi.closure import IN, OUT, ClosureScene, scene_pair_data, stable_substeps

    sc = ClosureScene.random(2000, **bench.CFG4_SCENES)
    parents, children = sc.particles()
    assert sc.shape == (96, 48, 48) and len(children) == 8 * len(parents)
    assert all(4.0 <= a <= 10.0 for a in sc.semi_axes) and 2.0e2 <= sc.youngs_modulus <= 2.0e4
    assert sc.substeps == stable_substeps(sc.youngs_modulus, sc.poisson_ratio, sc.density)
    fine = sc.simulation(True)
    assert (fine.lbm.nx, fine.lbm.ny, fine.lbm.nz) == (192, 96, 96)
    assert fine.mpm.num_particles == len(children)
    del fine
    X, T = scene_pair_data(sc, 10, every=4, warmup=2)  # coarse steps 2 and 6
    print(f"{len(parents)} coarse particles, {sc.substeps} substeps: {X.shape[0]} rows")
    assert X.shape == (2 * len(parents), IN) and T.shape == (2 * len(parents), OUT)
    assert np.all(np.isfinite(X)) and np.all(np.isfinite(T))
    support = X[:, 16]
    assert np.all((support > 0.0) & (support <= 1.0 + 1e-5))
    assert np.abs(T).max() > 0.0


def test_pair_runs_transfer_the_same_momentum():
    """The premise of the targets: over a coarse step, the refined run's
    fluid-on-solid force (children summed, in coarse units) and the coarse
    run's agree in total to within the resolution error, while the
    per-particle differences the closure learns are the larger part."""
    import bench
    from fsi.closure import ClosureScene, run_observables

    sc = ClosureScene.random(2001, **bench.CFG4_SCENES)
    bodies = sc.bodies()
    coarse = run_observables(sc.simulation(False), 12, bodies, force_every=12)
    fine = run_observables(sc.simulation(True), 12, bodies, refined=True, force_every=12)
    fc, ff = coarse["particle_forces"][-1], fine["particle_forces"][-1]
    total_rel = np.abs(fc.sum(0) - ff.sum(0)).max() / np.abs(ff.sum(0)).max()
    rms_rel = np.sqrt(((fc - ff) ** 2).sum() / (ff ** 2).sum())
    drag_rel = abs(coarse["drag_x"][-1] / fine["drag_x"][-1] - 1.0)
    print(f"total force {total_rel:.3e}, per-particle RMS {rms_rel:.3e}, drag {drag_rel:.3e}")
    assert np.all(np.isfinite(fc)) and np.all(np.isfinite(ff))
    assert drag_rel < 0.3 and total_rel < 0.3
    assert rms_rel > total_rel


def _large_scene():
    from fsi.closure import PackScene, stable_substeps

    return PackScene(shape=(192, 96, 96), count=64, radii=(2.5, 7.0), poisson_ratio=0.3, inflow=0.04, margin=4.0,
                     gap=2.0, seed=5, body_modulus=(5.0e2, 5.0e3), substeps=stable_substeps(5.0e3, 0.3, 1.2))


@pytest.mark.parametrize("refined,storage", [(True, "fp32"), (False, "fp16")])
def test_large_scene_at_the_configuration_size(refined, storage):
    import torch

    sc = _large_scene()
    E = sc.body_moduli()
    assert len(sc.bodies()) == 64 and np.all((E >= 5.0e2) & (E <= 5.0e3))
    sim = sc.simulation(refined, storage)
    f = 2 if refined else 1
    assert (sim.lbm.nx, sim.lbm.ny, sim.lbm.nz) == (192 * f, 96 * f, 96 * f)
    n = sim.mpm.num_particles
    assert n == f ** 3 * len(sc.particles()[0])
    # the material plane holds each sphere's Lame parameters (2 mu, lambda)
    two_mu = sim.mpm.material[:n].cpu().numpy().astype(np.float64)
    expect = np.repeat(E, np.asarray([b - a for a, b in sc.bodies()]) * f ** 3) / (1.0 + 0.3)
    assert np.allclose(np.sort(two_mu), np.sort(expect), rtol=1e-6)
    sim.coupler.record_step_force()
    x0 = sim.mpm.positions_numpy().astype(np.float64)
    for _ in range(2):
        sim.step()
    d = sim.diagnostics()
    fp = sim.coupler.step_force_numpy().astype(np.float64).sum(0)
    ff = sim.coupler.coupling_force_numpy().astype(np.float64)[sim.lbm.solid_numpy() == 0].sum(0)
    u = sim.lbm.velocity_numpy()
    x = sim.mpm.positions_numpy().astype(np.float64)
    print(f"{'refined' if refined else 'coarse'} {storage}: {n} particles, particles {fp}, fluid {ff}, "
          f"|u| max {np.abs(u).max():.4f}")
    assert d["coupling_unsupported_particle_count"] == 0
    assert np.abs(fp + ff).max() <= 1e-4 * np.abs(fp).max()
    assert np.all(np.isfinite(u)) and np.all(np.isfinite(x))
    assert 0.03 < np.abs(u).max() < 0.2 and np.abs(x - x0).max() < 2.0 * f
    del sim
    torch.cuda.empty_cache()
