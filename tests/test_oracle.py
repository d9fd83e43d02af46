"""Pins of the CPU oracle (no GPU).

The oracle is only trusted after it reproduces every committed reference
fixture (data/reference/*.json) and a set of closed forms / invariants chosen
so that a plausible restatement mistake (dropped term, wrong sign, wrong
index, floor instead of truncation) fails one of them.
"""

import numpy as np
import pytest

import cases
from oracle import pyoracle as po
from parity import porous_mask


@pytest.mark.parametrize("name", list(cases.FIXTURE_CASES))
def test_fp32_oracle_reproduces_reference_fixture(name):
    metrics = cases.fixture_metrics(cases.oracle_api("f32"), name)
    assert cases.compare_fixture(name, metrics) == []


def test_fixture_values_reproduced_bit_for_bit_where_quantised():
    """Fixture values that are exact FP32 quantities come out identically."""
    grav = cases.case_mpm_gravity(cases.oracle_api("f32"))
    assert grav["center_of_mass_y_delta"] == -1.9073486328125e-06
    drift = cases.case_coupled_drift(cases.oracle_api("f32"))
    assert drift["enabled_dx"] == 9.584426879882812e-05
    lim = cases.case_force_limit(cases.oracle_api("f32"))
    assert lim["particle_force_norm"] == 0.05000000074505806
    ib = cases.case_ib_drag(cases.oracle_api("f32"))
    assert ib["ib_total_force_x"] == -0.00032500000088475645
    assert ib["ib_active_cell_count"] == 19


def test_fp64_force_response_closed_form():
    """Guo forcing as written (lbm3d.py:276-283) gives a momentum source F/3 per
    step; with the +F/2 velocity shift, u_n = n F/3 + F/2 on a uniform periodic box."""
    m = cases.case_lbm_force_response(cases.oracle_api("f64"))
    F = 1.0e-6
    assert m["mean_ux_growth"] == pytest.approx(20 * F / 3 + F / 2, rel=1e-9)


def test_fp64_ib_drag_closed_form():
    """One particle of volume 4 at a node: phi = 4 * wx*wy*wz with per-axis
    weights (1/8, 3/4, 1/8); clamp at 1; 19 cells reach phi >= 0.01; the clipped
    sum of -drag*phi*rho*u over them is -0.1 * 3.25 * 1e-3."""
    m = cases.case_ib_drag(cases.oracle_api("f64"))
    assert m["ib_active_cell_count"] == 19
    assert m["ib_total_force_x"] == pytest.approx(-3.25e-4, rel=1e-12)


def test_truncating_cell_key_partial_support():
    """x = -0.75: int(-1.25) = -1 (truncation), fx = 0.25 -> valid weight of the
    two in-grid nodes 0.1875 + 0.03125; floor would give 0.28125."""
    m = cases.case_boundary_support(cases.oracle_api("f64"))
    assert m["min_particle_valid_weight"] == 0.21875
    assert m["partial_support_particle_count"] == 1


def test_contact_damping_exact():
    m = cases.case_contact(cases.oracle_api("f64"))
    assert m["contact_candidate_count"] == 1
    # the initial speed enters through float32 (the reference's field type)
    assert m["particle_speed_delta"] == pytest.approx(-0.25 * float(np.float32(0.1)), abs=1e-15)


def test_mass_budget_periodic_with_solids():
    """Streaming and bounce-back move mass exactly; the only source is the
    conserved-moment Guo term G_0 = -2/3 u.F per fluid cell and step, so the
    mass change of each step equals sum(-2/3 u.F) over the fluid cells."""
    cfg = cases.periodic_lbm_config(nx=12, ny=10, nz=8, force=(1e-5, 2e-6, 0.0))
    solid = porous_mask((12, 10, 8), seed=1, porosity=0.75, rmin=1.0, rmax=2.0)
    s = po.OracleLBM(cfg, "f64")
    s.initialize(solid)
    fluid = solid == 0
    for _ in range(30):
        m0 = s.rho[fluid].sum()
        predicted = (-2.0 / 3.0 * (s.v * s.force).sum(-1))[fluid].sum()
        s.step()
        assert s.rho[fluid].sum() - m0 == pytest.approx(predicted, rel=1e-6, abs=1e-11)
    zero = po.OracleLBM(cases.periodic_lbm_config(nx=12, ny=10, nz=8), "f64")
    zero.initialize(solid)
    m0 = zero.total_mass()
    for _ in range(30):
        zero.step()
    assert abs(zero.total_mass() - m0) / m0 < 1e-13


def test_rigid_translation_keeps_F_identity_and_C_zero():
    cfg = cases.small_mpm_config(gravity=(0.0, 0.0, 0.0), dt=1e-2)
    s = po.OracleMPM(cfg, "f64")
    s.initialize_particles_box((6.0, 6.0, 6.0), (10.0, 10.0, 10.0), spacing=0.5,
                               initial_velocity=(0.1, -0.05, 0.02))
    for _ in range(5):
        s.substep()
    n = s.num_particles
    assert np.abs(s.F[:n] - np.eye(3)).max() < 1e-12
    assert np.abs(s.C[:n]).max() < 1e-10
    assert np.allclose(s.v[:n], [0.1, -0.05, 0.02], atol=1e-12)


def test_polar_rotation_properties():
    rng = np.random.default_rng(3)
    for i in range(200):
        F = np.eye(3) + 0.4 * rng.standard_normal((3, 3))
        if i % 5 == 0:
            F[0] *= -1.0  # inverted
        R = po.polar_rotation(F, "f64")
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-12)
        assert np.linalg.det(R) == pytest.approx(1.0, abs=1e-12)
        S = R.T @ F
        assert np.allclose(S, S.T, atol=1e-11)
        if np.linalg.det(F) > 0:
            # polar factor: S positive definite; matches the SVD U V^T
            U, _, Vt = np.linalg.svd(F)
            assert np.allclose(R, U @ Vt, atol=1e-10)
            assert np.all(np.linalg.eigvalsh(S) > 0)


def test_two_relaxation_time_form_of_the_collision():
    """The lattice kernels collide in the two-relaxation-time form
    f* = F - s_v (F - g)^+ - s_o (F - g)^- + w e.F + c_v w ((e.u)(e.F) - u.F)
    - (1 - c_v) (2/3) w u.F, g = w (rho_F + 3 e.j + 4.5 (e.u)^2 - 1.5 u.u).
    It must equal the reference's matrix form m = M F; m -= S (m - meq);
    m += (1 - S/2) G; f = M^-1 m (fsi/lbm3d.py:270-333) for any F, u, force
    (the collision's rho and u need not be F's own), at both rates."""
    M = po.MRT_MATRIX
    Minv = np.linalg.inv(M)
    E = po.E.astype(np.float64)
    w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    rng = np.random.default_rng(5)

    def meq(rho, u):
        ux, uy, uz = u
        m = np.zeros(19)
        m[0], m[1], m[3], m[5], m[7] = rho, u @ u, ux, uy, uz
        m[9], m[11] = 2 * ux * ux - uy * uy - uz * uz, uy * uy - uz * uz
        m[13], m[14], m[15] = ux * uy, uy * uz, ux * uz
        return m

    for tau in (0.51, 0.55, 0.8, 1.7):
        S = po.relaxation_rates((tau - 0.5) / 3.0)
        s_v, s_o = S[1], S[4]
        c_v = 1 - s_v / 2
        for _ in range(20):
            F = w * (1 + rng.normal(0, 0.05, 19))
            rho, u, fo = rng.uniform(0.9, 1.1), rng.normal(0, 0.05, 3), rng.normal(0, 1e-3, 3)
            G = M @ (w * ((E - u) @ fo + (E @ u) * (E @ fo)))
            m = M @ F
            ref = Minv @ (m - S * (m - meq(rho, u)) + (1 - 0.5 * S) * G)
            g = w * (F.sum() + 3 * E @ (E.T @ F) + 4.5 * (E @ u) ** 2 - 1.5 * u @ u)
            h = F - g
            hp, hm = 0.5 * (h + h[po.LR]), 0.5 * (h - h[po.LR])
            trt = (F - s_v * hp - s_o * hm + w * (E @ fo) + c_v * w * ((E @ u) * (E @ fo) - u @ fo)
                   - (1 - c_v) * (2 / 3) * w * (u @ fo))
            assert np.abs(trt - ref).max() < 1e-15
