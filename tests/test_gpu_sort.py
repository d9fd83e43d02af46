"""Cell-key counting sort of the MPM particles (fsi_mpm_sort).

The permutation must equal a stable CPU sort (numpy argsort, kind="stable")
of keys computed on the host with the same float32 arithmetic; the accessors
must return the particles in input order whatever the slot order; and runs
with and without sorting must agree to the rounding of the grid sums.
"""

import numpy as np
import pytest

from parity import close, coupled_scene, sim_fields

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("gpu_available")]


def host_keys(pos, active, n3, dx=1.0):
    """Blocked cell keys of fsi_mpm_sort restated on the host: float32
    x * inv_dx, then - 0.5, truncated; clamped to [-1, n]; shifted by one;
    4x4x4 blocks, row-major with an even number of blocks in z; inactive
    particles take the last key."""
    inv = np.float32(1.0 / dx)
    gp = (pos.astype(np.float32) * inv).astype(np.float32)
    s = (gp - np.float32(0.5)).astype(np.float32)
    base = np.trunc(s).astype(np.int64)
    n3 = np.asarray(n3)
    c = np.clip(base, -1, n3) + 1
    pb = (n3 + 2 + 3) // 4
    pb[2] += pb[2] % 2  # an even number of z blocks
    blk = ((c[:, 0] >> 2) * pb[1] + (c[:, 1] >> 2)) * pb[2] + (c[:, 2] >> 2)
    key = blk * 64 + ((c[:, 0] & 3) * 4 + (c[:, 1] & 3)) * 4 + (c[:, 2] & 3)
    nkeys = int(pb.prod()) * 64
    return np.where(active != 0, key, nkeys).astype(np.uint32)


def _solver(n3, count):
    import fsi
    from fsi.config import MPMConfig

    cfg = MPMConfig(nx=n3[0], ny=n3[1], nz=n3[2], max_particles=count + 7, dt=1e-3, gravity=(0.0, 0.0, 0.0))
    return fsi.MPMSolver3D(cfg)


def _sort(s, n):
    import torch

    keys = torch.zeros(n, dtype=torch.int32, device=s.device)
    perm = torch.zeros(n, dtype=torch.int32, device=s.device)
    s.sort_particles(keys_out=keys, perm_out=perm)
    return keys.cpu().numpy().view(np.uint32), perm.cpu().numpy()


def test_sort_is_a_stable_sort_of_the_host_keys():
    import torch

    rng = np.random.default_rng(5)
    n3 = (37, 29, 23)
    n = 50_000
    pos = rng.uniform(-2.0, 40.0, size=(n, 3)).astype(np.float32)
    # exact cell faces and half-cells, where truncation and rounding matter
    pos[:500] = rng.integers(0, 20, size=(500, 3)).astype(np.float32) + np.float32(0.5)
    pos[500:1000] = rng.integers(0, 20, size=(500, 3)).astype(np.float32)
    pos[1000] = (-0.75, 4.0, 4.0)
    s = _solver(n3, n)
    vel = rng.normal(size=(n, 3)).astype(np.float32)
    s.initialize_particles_from_numpy(pos, vel)
    inactive = rng.choice(n, 300, replace=False)
    s.active[torch.as_tensor(inactive, device=s.device)] = 0
    active = s.active_numpy()
    ref_keys = host_keys(s.positions_numpy(), active, n3)
    expected = np.argsort(ref_keys, kind="stable")

    keys, perm = _sort(s, n)
    assert np.array_equal(perm, expected)
    assert np.array_equal(keys, ref_keys[expected])
    assert np.all(np.diff(keys.astype(np.int64)) >= 0)
    # the accessors still answer in input order
    assert np.array_equal(s.positions_numpy(), pos)
    assert np.array_equal(s.velocities_numpy(), vel)
    assert np.array_equal(s.active_numpy(), active)
    assert np.array_equal(s.ids[:n].cpu().numpy(), expected)

    # a second sort from the sorted state keeps everything (ties by slot)
    keys2, perm2 = _sort(s, n)
    assert np.array_equal(perm2, np.arange(n))
    assert np.array_equal(keys2, keys)
    assert np.array_equal(s.positions_numpy(), pos)


def test_crowded_cell_uses_the_long_segment_path():
    n = 5000
    s = _solver((16, 16, 16), n + 10)
    pos = np.full((n + 10, 3), 7.3, np.float32)
    pos[n:] = 3.1  # a second, smaller cell, sorted first
    s.initialize_particles_from_numpy(pos)
    keys, perm = _sort(s, n + 10)
    assert np.array_equal(perm, np.concatenate([np.arange(n, n + 10), np.arange(n)]))
    assert np.array_equal(s.positions_numpy(), pos)


def test_mpm_substeps_after_a_sort_match_unsorted():
    import fsi
    from fsi.config import MPMConfig

    rng = np.random.default_rng(2)
    cfg = MPMConfig(nx=24, ny=24, nz=24, max_particles=6000, dt=2e-3, youngs_modulus=500.0,
                    gravity=(0.0, -1e-3, 0.0), boundary_width=2)
    pos = rng.uniform(6.0, 18.0, size=(6000, 3)).astype(np.float32)
    rng.shuffle(pos)
    vel = rng.normal(0, 0.05, size=(6000, 3)).astype(np.float32)
    out = {}
    for sort in (False, True):
        s = fsi.MPMSolver3D(cfg, precision="fp64")
        s.initialize_particles_from_numpy(pos, vel, particle_mass=0.2, particle_volume=0.125)
        for k in range(20):
            if sort and k % 5 == 0:
                s.sort_particles()
            s.substep()
        out[sort] = (s.positions_numpy(), s.velocities_numpy(), s.deformation_gradients_numpy())
    for a, b in zip(out[True], out[False]):
        ok, err, _ = close(a, b, 1e-11, 1e-13)
        assert ok, err


@pytest.mark.parametrize("storage", ["fp64", "fp32"])
def test_coupled_steps_with_and_without_sorting(storage):
    import fsi

    fields = {}
    for interval in (0, 1):
        cfg, solid, pos, vel = coupled_scene(storage=storage, ib=True, contact=True)
        sim = fsi.FSISimulation(cfg)
        sim.initialize_lbm(solid)
        sim.coupler.sort_interval = interval
        # shuffled input order: sorting really permutes
        order = np.random.default_rng(1).permutation(len(pos))
        sim.initialize_mpm_from_numpy(pos[order], vel[order], particle_mass=0.15, particle_volume=0.125)
        for _ in range(4):
            sim.advance()
        fields[interval] = sim_fields(sim)
    rel = 1e-10 if storage == "fp64" else 2e-5
    for key, a in fields[1].items():
        b = fields[0][key]
        if a.dtype.kind in "iu":
            assert np.array_equal(a, b), key
        else:
            ok, err, _ = close(a, b, rel, 1e-7 if storage == "fp32" else 1e-13)
            assert ok, (key, err)
