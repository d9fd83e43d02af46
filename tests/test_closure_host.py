"""Host logic of the closure's data pairs and accuracy measures (CPU): the
refinement of particles, the scene geometry, the refined configuration's
scaling and the error measures, against values fixed by construction."""

import numpy as np
import pytest

from fsi.closure import ClosureScene, PackScene, observable_errors, pair_config, refine_particles


def test_children_fill_the_parent_volume():
    parents = np.array([[3.25, 4.75, 5.25], [10.0, 2.0, 7.5]], np.float32)
    kids = refine_particles(parents).reshape(2, 8, 3).astype(np.float64)
    # the children's centroid is the parent position in fine units, and they
    # sit on the corners of a cube of side 0.5 fine cells around it
    np.testing.assert_allclose(kids.mean(1), 2.0 * parents, atol=1e-6)
    for c in kids:
        d = c - c.mean(0)
        assert np.allclose(np.abs(d), 0.25) and len({tuple(np.sign(r)) for r in d}) == 8


def test_refined_config_keeps_the_physical_problem():
    c = pair_config((24, 12, 12), False, 100, 1.0e3, 0.3, 0.03)
    f = pair_config((24, 12, 12), True, 800, 1.0e3, 0.3, 0.03)
    assert (f.lbm.nx, f.lbm.ny, f.lbm.nz) == (48, 24, 24)
    assert f.lbm.viscosity == pytest.approx(2 * c.lbm.viscosity)      # same Reynolds number
    assert f.coupling.gamma == pytest.approx(c.coupling.gamma / 2)     # same rate per physical time
    assert f.mpm.boundary_width == 2 * c.mpm.boundary_width            # same physical layer
    assert f.mpm.youngs_modulus == c.mpm.youngs_modulus and f.mpm.density == c.mpm.density
    assert f.lbm.x_left.velocity == c.lbm.x_left.velocity              # same lattice velocity


def test_pack_scene_bodies_and_placement():
    sc = PackScene.random(501)
    parents, children = sc.particles()
    bodies = sc.bodies()
    assert len(bodies) == sc.count and bodies[0][0] == 0 and bodies[-1][1] == len(parents)
    assert all(a < b for a, b in bodies) and all(bodies[i][1] == bodies[i + 1][0] for i in range(len(bodies) - 1))
    assert len(children) == 8 * len(parents)
    # packed into the upstream part of the channel, clear of the walls
    assert parents[:, 0].max() < sc.upstream * sc.shape[0]
    assert parents[:, 1:].min() > sc.margin - 0.5 and parents[:, 1].max() < sc.shape[1] - sc.margin + 0.5
    # every body is one connected ball: its particles lie within its radius bound
    for a, b in bodies:
        p = parents[a:b].astype(np.float64)
        assert np.linalg.norm(p - p.mean(0), axis=1).max() <= sc.radii[1] + 0.5
    assert ClosureScene.random(3).bodies() == [(0, len(ClosureScene.random(3).particles()[0]))]


def test_error_measures():
    fine = {"drag_x": np.array([4.0, 2.0, 1.0, 1.0]), "displacement": np.array([[1.0, 0.0, 0.0], [0.0, 2.0, 0.0]])}
    run = {"drag_x": np.array([4.4, 2.2, 1.1, 1.1]), "displacement": np.array([[1.0, 0.1, 0.0], [0.0, 2.0, 0.0]])}
    e = observable_errors(run, fine)
    assert e["drag_rel_err"] == pytest.approx(0.1) and e["drag_rel_err_second_half"] == pytest.approx(0.1)
    # vector error: |(0, 0.1, 0)| / |(1, 0, 0)| for body 0, 0 for body 1
    assert e["displacement_rel_err"] == pytest.approx(0.05) and e["displacement_rel_err_max"] == pytest.approx(0.1)


def test_stable_substeps_hold_the_p_wave_courant_number():
    """S = ceil(c_p / 0.3) with c_p = sqrt(E (1 - nu) / ((1 + nu)(1 - 2 nu)) / rho),
    at least 112: the 1e3 material of the coupled scenes needs 112 and the
    stiffest sphere of the suspension (5e3) 250, as their configurations
    state; the refined run of a pair keeps the count per lattice step."""
    from fsi.closure import stable_substeps

    assert stable_substeps(1.0e3, 0.3, 1.2) == 112
    assert stable_substeps(5.0e3, 0.3, 1.2) == 250
    assert stable_substeps(2.0e2, 0.2, 1.2) == 112  # soft material: the floor
    for E, nu in ((2.0e4, 0.4), (7.3e3, 0.25), (3.1e3, 0.33)):
        s = stable_substeps(E, nu, 1.2)
        c = np.sqrt(E * (1 - nu) / ((1 + nu) * (1 - 2 * nu)) / 1.2)
        assert c / s <= 0.3 < c / (s - 1)
    sc = ClosureScene.random(1013, shape=(96, 48, 48), axes=(4.0, 10.0), modulus=(2.0e2, 2.0e4), substeps="stable")
    assert sc.substeps == stable_substeps(sc.youngs_modulus, sc.poisson_ratio, sc.density)
    assert sc.config(refined=True).coupling.mpm_substeps_per_lbm_step == sc.substeps


def test_pack_scene_per_body_moduli():
    """Per-sphere E log-uniform in the range, one value per body, and the
    configuration's substeps taken from the stiffest."""
    sc = PackScene(shape=(48, 24, 24), count=4, radii=(2.5, 4.0), seed=3, body_modulus=(5.0e2, 5.0e3))
    E = sc.body_moduli()
    assert len(E) == len(sc.bodies()) and np.all((E >= 5.0e2) & (E <= 5.0e3)) and len(set(E.tolist())) == len(E)
    assert sc.config().mpm.youngs_modulus == pytest.approx(E.max())
    assert PackScene(shape=(48, 24, 24), count=4, radii=(2.5, 4.0), seed=3).body_moduli() is None
