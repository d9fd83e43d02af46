"""Per-particle elastic material (MPMSolver3D.set_particle_material).

The reference has one global Young's modulus and Poisson ratio
(fsi/mpm3d.py:33-38); bodies of different stiffness in one scene are a
build-only extension, read by the P2G stress from a per-particle (2 mu,
lambda) plane. Its results are pinned to the global-material path, which
the oracle pins:

* a plane holding the config's own E and nu for every particle gives the
  run without a plane (FP64 build to 1e-12; the only difference is the
  order of float atomics);
* two well-separated bodies of different E in one run move as each body
  does alone in a run whose global E is its own (their stencils never
  share a node, so each body's grid sums are its own);
* the same through the coupled substeps (fused G2P+P2G launches over the
  cell-sorted work list, so the plane is also permuted by the sort), with a
  cube whose plane E differs from the config's.
"""

import numpy as np
import pytest

from cases import small_mpm_config
from parity import coupled_scene, jittered_cube, rel_inf, sim_fields

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("gpu_available")]

E1, E2, NU2 = 200.0, 2000.0, 0.35


def _bodies():
    """Two jittered cubes 6 cells apart, each squeezed along x and sheared."""
    a = jittered_cube((2.0, 4.0, 4.0), (6.0, 8.0, 8.0), 0.5, seed=3)
    b = jittered_cube((12.0, 6.0, 5.0), (16.0, 10.0, 9.0), 0.5, seed=4)
    vel = []
    for p in (a, b):
        c = p.mean(0)
        v = np.zeros_like(p)
        v[:, 0] = -0.4 * (p[:, 0] - c[0])
        v[:, 1] = 0.3 * (p[:, 2] - c[2])
        vel.append(v.astype(np.float32))
    return (a, vel[0]), (b, vel[1])


def _run(precision, pos, vel, E, nu=0.25, material=None, steps=60, sort=True):
    import fsi

    cfg = small_mpm_config(nx=24, youngs_modulus=E, poisson_ratio=nu, dt=0.005)
    s = fsi.MPMSolver3D(cfg, precision=precision)
    s.initialize_particles_from_numpy(pos, vel, particle_volume=0.125)
    if material is not None:
        s.set_particle_material(*material)
    if sort:
        s.sort_particles()
    for _ in range(steps):
        s.substep()
    return {"x": s.positions_numpy(), "v": s.velocities_numpy(), "F": s.deformation_gradients_numpy()}


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_plane_of_the_global_material_changes_nothing(precision):
    (a, va), (b, vb) = _bodies()
    pos, vel = np.concatenate([a, b]), np.concatenate([va, vb])
    plain = _run(precision, pos, vel, E2, nu=NU2)
    plane = _run(precision, pos, vel, E2, nu=NU2, material=(np.full(len(pos), E2), NU2))
    tol = 1e-12 if precision == "fp64" else 1e-5
    for k in plain:
        assert rel_inf(plane[k], plain[k]) <= tol, (k, rel_inf(plane[k], plain[k]))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_two_bodies_move_as_each_alone_with_its_own_material(precision):
    (a, va), (b, vb) = _bodies()
    pos, vel = np.concatenate([a, b]), np.concatenate([va, vb])
    E = np.concatenate([np.full(len(a), E1), np.full(len(b), E2)])
    nu = np.concatenate([np.full(len(a), 0.25), np.full(len(b), NU2)])
    both = _run(precision, pos, vel, 700.0, nu=0.3, material=(E, nu))
    alone_a = _run(precision, a, va, E1, nu=0.25)
    alone_b = _run(precision, b, vb, E2, nu=NU2)
    tol = 1e-12 if precision == "fp64" else 1e-5
    for k in both:
        ea = rel_inf(both[k][: len(a)], alone_a[k])
        eb = rel_inf(both[k][len(a):], alone_b[k])
        print(f"{precision} {k}: body a {ea:.2e}, body b {eb:.2e}")
        assert ea <= tol and eb <= tol, (k, ea, eb)
    # the materials differ enough to matter: body b under body a's material
    # ends elsewhere
    swapped = _run(precision, b, vb, E1, nu=0.25)
    assert rel_inf(swapped["F"], alone_b["F"]) > 1e-4


def test_coupled_substeps_with_a_material_plane():
    """FP64 coupled scene (fused substep launches over the sorted work list):
    the cube's plane at E = 600 reproduces the run whose config has E = 600."""
    import fsi

    def run(e_cfg, plane_e):
        cfg, solid, pos, vel = coupled_scene(storage="fp64", e_mod=e_cfg)
        sim = fsi.FSISimulation(cfg)
        sim.initialize_lbm(solid)
        sim.initialize_mpm_from_numpy(pos, vel, particle_volume=0.125)
        if plane_e is not None:
            sim.mpm.set_particle_material(np.full(len(pos), plane_e))
        for _ in range(6):
            sim.step()
        return sim_fields(sim)

    ref = run(600.0, None)
    got = run(400.0, 600.0)
    for k in ("rho", "u", "x", "v", "F", "pforce", "cforce"):
        e = rel_inf(got[k], ref[k])
        print(f"coupled {k}: {e:.2e}")
        assert e <= 1e-11, (k, e)
    assert np.array_equal(got["cmask"], ref["cmask"])


def test_material_arguments_are_checked():
    import fsi

    s = fsi.MPMSolver3D(small_mpm_config())
    with pytest.raises(RuntimeError, match="no particles"):
        s.set_particle_material(100.0)
    (a, va), _ = _bodies()
    s.initialize_particles_from_numpy(a, va)
    with pytest.raises(ValueError, match="youngs_modulus"):
        s.set_particle_material(-1.0)
    with pytest.raises(ValueError, match="poisson_ratio"):
        s.set_particle_material(100.0, 0.5)
    s.set_particle_material(100.0)
    assert s.material is not None
    s.initialize_particles_from_numpy(a, va)
    assert s.material is None
