"""FP64 device reductions behind the diagnostics (fsi_reduce_lattice,
fsi_reduce_particles) against FP64 numpy sums of the same fields copied to the
host, plus the diagnostics dict against the reference's float32 host
formulas (fsi/simulation.py:117-147)."""

import numpy as np
import pytest

from parity import coupled_scene

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("gpu_available")]


def _porous_lattice(n=96, steps=3):
    """96^3 (> 1024 workgroups x 256 lanes, so the grid-stride loop runs
    several rounds), random spheres, a little flow."""
    import fsi
    from fsi.config import BoundaryConfig, LBMConfig
    from fsi.scenes import porous_spheres

    faces = {f: BoundaryConfig("periodic") for f in ("x_left", "x_right", "y_left", "y_right", "z_left", "z_right")}
    cfg = LBMConfig(nx=n, ny=n, nz=n, viscosity=0.05, force=(1e-5, -2e-6, 3e-6),
                    initial_velocity=(0.01, 0.002, -0.003), **faces)
    mask, _ = porous_spheres((n, n, n), porosity=0.85, rmin=3.0, rmax=6.0, seed=3)
    s = fsi.LBMSolver3D(cfg)
    s.initialize(mask)
    for _ in range(steps):
        s.step()
    return s, mask


def test_lattice_sums_match_fp64_numpy():
    s, mask = _porous_lattice()
    sums = s.lattice_sums().cpu().numpy()
    rho = s.density_numpy().astype(np.float64)
    u = s.velocity_numpy().astype(np.float64)
    mass = rho[mask == 0].sum()
    umax2 = (u * u).sum(axis=-1).max()
    assert sums[0] == pytest.approx(mass, rel=1e-13)
    assert sums[1] == pytest.approx(umax2, rel=1e-14)
    assert np.all(sums[2:5] == 0.0)  # no coupling buffer given
    # the accessors keep the reference's float32 host formulas: equal to
    # float32 rounding
    assert s.total_mass() == pytest.approx(mass, rel=1e-6)
    assert s.max_velocity_norm() == pytest.approx(np.sqrt(umax2), rel=1e-6)
    # same grid, same order: bit-identical on a repeat
    assert np.array_equal(s.lattice_sums().cpu().numpy(), sums)


def test_particle_sums_match_fp64_numpy():
    import fsi
    from fsi.config import MPMConfig

    rng = np.random.default_rng(11)
    n = 300_000  # > 1024 x 256: several grid-stride rounds
    cfg = MPMConfig(nx=64, ny=64, nz=64, max_particles=n + 5, dt=1e-3, gravity=(0.0, 0.0, 0.0))
    s = fsi.MPMSolver3D(cfg)
    pos = rng.uniform(2.0, 60.0, size=(n, 3)).astype(np.float32)
    vel = rng.normal(0.0, 0.1, size=(n, 3)).astype(np.float32)
    s.initialize_particles_from_numpy(pos, vel, particle_mass=0.5, particle_volume=0.25)
    mass = s.masses_numpy().astype(np.float64)
    x = s.positions_numpy().astype(np.float64)
    v = s.velocities_numpy().astype(np.float64)
    sums = s.particle_sums().cpu().numpy()
    assert sums[0] == pytest.approx(mass.sum(), rel=1e-13)
    np.testing.assert_allclose(sums[1:4], (x * mass[:, None]).sum(axis=0), rtol=1e-13)
    assert sums[4] == pytest.approx((v * v).sum(axis=1).max(), rel=1e-14)
    assert np.all(sums[5:8] == 0.0)
    com = s._com_from(sums)
    assert com.dtype == np.float32
    np.testing.assert_allclose(com, (x * mass[:, None]).sum(axis=0) / mass.sum(), rtol=1e-7)
    # the reference's float32 host formula (the accessor) accumulates
    # ~1e-5 relative error over 3e5 particles; the FP64 sums do not
    np.testing.assert_allclose(s.center_of_mass(), com, rtol=1e-4)
    assert np.array_equal(s.particle_sums().cpu().numpy(), sums)


def test_particle_max_propagates_nan():
    import fsi
    from fsi.config import MPMConfig

    s = fsi.MPMSolver3D(MPMConfig(nx=16, ny=16, nz=16, max_particles=8, gravity=(0.0, 0.0, 0.0)))
    pos = np.full((3, 3), 8.0, np.float32)
    vel = np.zeros((3, 3), np.float32)
    vel[1, 2] = np.nan
    s.initialize_particles_from_numpy(pos, vel)
    assert np.isnan(s.particle_sums()[4].item())
    assert np.isnan(s.max_velocity_norm())


def test_diagnostics_match_host_formulas():
    import fsi

    cfg, solid, pos, vel = coupled_scene(ib=True, contact=True)
    sim = fsi.FSISimulation(cfg)
    sim.initialize_lbm(solid)
    sim.initialize_mpm_from_numpy(pos, vel, particle_mass=0.15, particle_volume=0.125)
    for _ in range(3):
        d = sim.step()
    assert len(d) == 18
    lbm, mpm, cpl = sim.lbm, sim.mpm, sim.coupler
    rho, u, sld = lbm.density_numpy(), lbm.velocity_numpy(), lbm.solid_numpy()
    x, m = mpm.positions_numpy(), mpm.masses_numpy()
    f = mpm.particle_forces_numpy()
    cf = cpl.coupling_force_numpy()
    w = cpl.particle_valid_weights_numpy()
    # the reference's float32 host reductions (fsi/lbm3d.py:449-456,
    # fsi/mpm3d.py:353-368, fsi/coupling.py:238-244, 263-270)
    assert d["lbm_total_mass"] == pytest.approx(float(rho[sld == 0].sum()), rel=1e-6)
    assert d["lbm_max_velocity_norm"] == pytest.approx(float(np.linalg.norm(u, axis=-1).max()), rel=1e-6)
    np.testing.assert_allclose(d["mpm_center_of_mass"], (x * m[:, None]).sum(axis=0) / m.sum(), rtol=1e-6)
    assert d["mpm_max_velocity_norm"] == pytest.approx(float(np.linalg.norm(mpm.velocities_numpy(), axis=1).max()),
                                                       rel=1e-6)
    scale = np.abs(f).sum(axis=0).max()
    np.testing.assert_allclose(d["total_particle_coupling_force"], f.astype(np.float64).sum(axis=0),
                               rtol=0, atol=1e-6 * scale)
    fluid = cf[sld == 0].astype(np.float64)
    np.testing.assert_allclose(d["total_fluid_coupling_force"], fluid.sum(axis=0), rtol=0,
                               atol=1e-6 * np.abs(fluid).sum(axis=0).max())
    cd = cpl.coupling_diagnostics()
    assert cd["min_particle_valid_weight"] == pytest.approx(float(w.min()), rel=1e-7)
    assert cd["mean_particle_valid_weight"] == pytest.approx(float(w.astype(np.float64).mean()), rel=1e-6)
    ctr = cpl.counters.cpu().numpy()
    assert d["coupling_contact_candidate_count"] == int(ctr[3])
    assert d["coupling_unsupported_particle_count"] == int(ctr[0])
    for key in ("step", "time", "mpm_particle_count", "coupling_effective_gamma"):
        assert key in d
