"""Seeded scene generator (fsi/scenes.py) on the CPU: determinism, the
porosity target, geometry checked by brute force on small grids, sphere
packs without overlap, per-body log-uniform moduli."""

import numpy as np
import pytest

from fsi import scenes


def test_porous_spheres_is_seeded_and_reaches_the_porosity():
    a, ia = scenes.porous_spheres((40, 32, 24), porosity=0.8, rmin=2.0, rmax=5.0, seed=3)
    b, _ = scenes.porous_spheres((40, 32, 24), porosity=0.8, rmin=2.0, rmax=5.0, seed=3)
    c, _ = scenes.porous_spheres((40, 32, 24), porosity=0.8, rmin=2.0, rmax=5.0, seed=4)
    assert a.dtype == np.int32 and a.shape == (40, 32, 24)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    solid = a.mean()
    assert ia["porosity"] == pytest.approx(1.0 - solid, abs=0.0)
    # spheres are added until the solid fraction reaches 0.2; the last one
    # overshoots by at most its own volume
    vmax = 4.0 / 3.0 * np.pi * 5.5 ** 3 / a.size
    assert 0.2 <= solid <= 0.2 + vmax


@pytest.mark.parametrize("periodic", [True, False])
def test_porous_spheres_mask_is_the_union_of_its_spheres(periodic):
    shape = (20, 18, 16)
    mask, info = scenes.porous_spheres(shape, porosity=0.85, rmin=2.0, rmax=4.0, seed=1, periodic=periodic)
    centres = np.stack(np.meshgrid(*[np.arange(n) + 0.5 for n in shape], indexing="ij"), -1)
    expect = np.zeros(shape, bool)
    for cx, cy, cz, r in info["spheres"]:
        d = centres - np.array([cx, cy, cz])
        if periodic:
            d -= np.round(d / np.array(shape)) * np.array(shape)
        expect |= (d ** 2).sum(-1) <= r * r
    assert np.array_equal(mask.astype(bool), expect)


def test_channel_walls():
    w = scenes.channel_walls((6, 5, 4))
    assert w[:, 0, :].all() and w[:, -1, :].all() and w[:, :, 0].all() and w[:, :, -1].all()
    assert not w[:, 1:-1, 1:-1].any()
    assert not scenes.channel_walls((6, 5, 4), y=False)[:, 1:-1, 1:-1].any()


def test_jittered_box_lattice_and_jitter():
    p = scenes.jittered_box((1.0, 2.0, 3.0), (3.0, 5.0, 4.0), 0.5, seed=7, jitter=0.5)
    assert p.dtype == np.float32 and p.shape == (4 * 6 * 2, 3)
    q = scenes.jittered_box((1.0, 2.0, 3.0), (3.0, 5.0, 4.0), 0.5, seed=7, jitter=0.5)
    assert np.array_equal(p, q)
    lattice = scenes.jittered_box((1.0, 2.0, 3.0), (3.0, 5.0, 4.0), 0.5, seed=7, jitter=0.0)
    off = np.abs(p - lattice)
    assert off.max() <= 0.125 + 1e-6 and off.max() > 0.05
    assert np.allclose(lattice[0], [1.25, 2.25, 3.25])


def test_rotations_and_ellipsoids():
    rng = np.random.default_rng(0)
    for _ in range(20):
        R = scenes.random_rotation(rng)
        assert np.allclose(R.T @ R, np.eye(3), atol=1e-12) and np.linalg.det(R) == pytest.approx(1.0)
    R = scenes.random_rotation(np.random.default_rng(5))
    c, semi = np.array([10.0, 9.0, 8.0]), np.array([4.0, 3.0, 2.0])
    p = scenes.jittered_ellipsoid(c, semi, 0.25, rotation=R, seed=2, jitter=0.0).astype(np.float64)
    local = (p - c) @ R
    assert (((local / semi) ** 2).sum(1) <= 1.0 + 1e-6).all()
    # lattice points fill the volume at one per spacing^3
    assert len(p) * 0.25 ** 3 == pytest.approx(4.0 / 3.0 * np.pi * semi.prod(), rel=0.03)


def test_sphere_pack_places_disjoint_spheres():
    shape = (64, 32, 32)
    pos, info = scenes.sphere_pack(shape, 6, 3.0, 5.0, spacing=0.5, seed=2)
    pos2, _ = scenes.sphere_pack(shape, 6, 3.0, 5.0, spacing=0.5, seed=2)
    assert np.array_equal(pos, pos2)
    c, r = info["centres"], info["radii"]
    assert len(r) == 6 and (r >= 3.0).all() and (r <= 5.0).all()
    for i in range(6):
        assert (c[i] - r[i] - 3.0 >= 0).all() and (c[i] + r[i] + 3.0 <= np.array(shape)).all()
        for j in range(i):
            assert np.linalg.norm(c[i] - c[j]) >= r[i] + r[j] + 1.0
    assert sum(info["counts"]) == len(pos)
    start = 0
    for i, n in enumerate(info["counts"]):
        body = pos[start:start + n].astype(np.float64)
        start += n
        assert (np.linalg.norm(body - c[i], axis=1) <= r[i] + 0.125 * np.sqrt(3) + 1e-5).all()
    with pytest.raises(ValueError, match="could not place"):
        scenes.sphere_pack((16, 16, 16), 20, 3.0, 5.0, max_tries=200)


def test_body_moduli_are_log_uniform_per_body():
    per_particle, per_body = scenes.body_moduli([3, 0, 2], 5.0e2, 5.0e3, seed=0)
    assert np.array_equal(per_particle, np.repeat(per_body, [3, 0, 2]))
    again, _ = scenes.body_moduli([3, 0, 2], 5.0e2, 5.0e3, seed=0)
    assert np.array_equal(per_particle, again)
    _, many = scenes.body_moduli(np.ones(20000, int), 5.0e2, 5.0e3, seed=1)
    assert many.min() >= 5.0e2 and many.max() <= 5.0e3
    logs = np.log(many)
    # uniform in log: mean at the midpoint, variance (log 10)^2 / 12
    assert logs.mean() == pytest.approx(0.5 * (np.log(5.0e2) + np.log(5.0e3)), abs=0.02)
    assert logs.var() == pytest.approx(np.log(10.0) ** 2 / 12.0, rel=0.03)
