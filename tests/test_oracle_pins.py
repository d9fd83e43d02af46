"""Mathematical pins of the oracle functions that no reference fixture covers.

Each test drives one oracle function (``oracle/fsi_oracle.cpp`` through
``oracle/pyoracle.py``) on a state for which mathematics fixes the answer,
chosen so that a plausible restatement mistake fails it:

* MRT collision (fsi/lbm3d.py:141-168, 285-333): a perturbation of one
  non-conserved moment relaxes by exactly (1 - S_k) and leaves the other 18
  moments alone; the equilibrium moments are a fixed point; a shear wave
  decays at nu k^2 with nu = (tau - 1/2)/3; a body-force channel between
  half-way bounce-back walls reaches the Poiseuille profile of force F/3.
* x-face boundaries (fsi/lbm3d.py:363-407): the velocity face's sequential
  in-place update equals its closed net form and keeps equilibrium fixed; the
  pressure face takes the velocity of the inner neighbour when that cell is
  solid.
* P2G stress and APIC (fsi/mpm3d.py:235-266, 291-318): the grid momentum of
  one particle recovers the fixed-corotated stress of an independent polar
  decomposition through the B-spline's second moment; an affine velocity
  field with C = A survives P2G -> grid update -> G2P exactly.
* Grid walls and position clamp (fsi/mpm3d.py:269-288, 321-328).
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import scipy.linalg

import cases
from oracle import pyoracle as po
from parity import W

# D3Q19 velocities (fsi/lbm3d.py:221-239)
E = po.E.astype(np.float64)
M = po.MRT_MATRIX
CONSERVED = (0, 3, 5, 7)


def meq(rho, u):
    """Equilibrium moments of fsi/lbm3d.py:285-312 (no density factor)."""
    ux, uy, uz = u
    m = np.zeros(19)
    m[0] = rho
    m[1] = ux * ux + uy * uy + uz * uz
    m[3], m[5], m[7] = ux, uy, uz
    m[9] = 2 * ux * ux - uy * uy - uz * uz
    m[11] = uy * uy - uz * uz
    m[13], m[14], m[15] = ux * uy, uy * uz, ux * uz
    return m


def feq(rho, u):
    eu = E @ np.asarray(u, np.float64)
    return W * rho * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * np.dot(u, u))


def rates(tau):
    s_v = 1.0 / tau
    s_o = 8.0 * (2.0 - s_v) / (8.0 - s_v)
    S = np.full(19, s_v)
    S[list(CONSERVED)] = 0.0
    S[[4, 6, 8, 16, 17, 18]] = s_o
    return S


def uniform_lattice(F_cell, u, n=(5, 5, 5), viscosity=0.1):
    """Periodic lattice holding the same populations in every cell, with the
    macroscopic rho/u the collision reads set to (1, u)."""
    cfg = cases.periodic_lbm_config(nx=n[0], ny=n[1], nz=n[2], viscosity=viscosity)
    s = po.OracleLBM(cfg, "f64")
    s.initialize()
    s.F[...] = F_cell
    s.f[...] = F_cell
    s.rho[...] = 1.0
    s.v[...] = u
    return s


@pytest.mark.parametrize("u", [(0.0, 0.0, 0.0), (0.03, -0.02, 0.015)])
def test_mrt_relaxes_each_moment_at_its_own_rate(u):
    """Uniform periodic state, zero force: streaming is the identity, so one
    step is one collision. Start from the equilibrium moments with moment k
    moved by delta: afterwards m_k - meq_k = (1 - S_k) delta and every other
    moment equals its equilibrium (conserved ones untouched)."""
    tau = 0.8
    S = rates(tau)
    invM = np.linalg.inv(M)
    delta = 1e-3
    for k in range(19):
        m0 = meq(1.0, u)
        m0[k] += delta
        s = uniform_lattice(invM @ m0, u, viscosity=(tau - 0.5) / 3.0)
        s.step()
        m1 = M @ s.F[2, 3, 1]
        expect = meq(1.0, u)
        expect[k] += (1.0 - S[k]) * delta
        assert np.abs(m1 - expect).max() < 1e-14, (k, m1 - expect)
        assert np.abs(s.F - s.F[0, 0, 0]).max() == 0.0  # still uniform


def test_mrt_equilibrium_moments_are_a_fixed_point():
    rng = np.random.default_rng(4)
    invM = np.linalg.inv(M)
    for _ in range(5):
        u = rng.uniform(-0.05, 0.05, 3)
        s = uniform_lattice(invM @ meq(1.0, u), u, viscosity=0.07)
        F0 = s.F[0, 0, 0].copy()
        s.step()
        assert np.abs(s.F[1, 2, 3] - F0).max() < 1e-15
        assert np.abs(s.v[0, 0, 0] - u).max() < 1e-15
        assert s.rho[0, 0, 0] == pytest.approx(1.0, abs=1e-15)


def test_shear_wave_decays_at_the_lattice_viscosity():
    """u_x = U sin(k y) on a periodic lattice decays as exp(-nu k^2 t) with
    nu = (tau - 1/2)/3 (the rate of the stress moments 9-15); relaxing them at
    the other rate s_o would give nu = 0.21 instead of 0.1."""
    ny, nu, U = 32, 0.1, 1e-4
    kw = 2.0 * np.pi / ny
    cfg = cases.periodic_lbm_config(nx=5, ny=ny, nz=5, viscosity=nu)
    s = po.OracleLBM(cfg, "f64")
    s.initialize()
    invM = np.linalg.inv(M)
    y = np.arange(ny)
    for j in range(ny):
        u = (U * np.sin(kw * y[j]), 0.0, 0.0)
        s.F[:, j] = invM @ meq(1.0, u)
        s.v[:, j] = u
    s.rho[...] = 1.0
    basis = np.sin(kw * y)

    def amplitude():
        return float((s.v[2, :, 2, 0] * basis).sum() / (basis * basis).sum())

    for _ in range(20):  # let the start-up transient of the odd moments pass
        s.step()
    a0, t = amplitude(), 400
    for _ in range(t):
        s.step()
    rate = -np.log(amplitude() / a0) / t
    assert rate == pytest.approx(nu * kw * kw, rel=2e-3), rate / (nu * kw * kw)
    assert np.abs(s.v[..., 1:]).max() < 1e-12


def test_poiseuille_channel_with_body_force_a_third():
    """Body force F_x between solid rows (half-way bounce-back). The
    reference's Guo term without the 1/c_s^2 factors is a momentum source
    F/3 per step (fsi/lbm3d.py:276-283), so the second-order momentum
    j + (F/3)/2 (j = sum e f) follows the Poiseuille profile
    F/3 / (2 nu) y (H - y), y measured from the walls half-way between the
    last fluid and the first solid row (exact for the relaxation pair
    s_o = 8 (2 - s_v)/(8 - s_v)). The reported velocity adds the full F/2
    (fsi/lbm3d.py:420-422): rho u = j + F/2, so rho u - F/3 is the profile.
    A standard Guo source (factor 3) would triple it; a shift of F/6 or no
    shift in the macro step would move it by F/3 or F/2 (~1e-3 relative)."""
    H, nu, Fx = 12, 0.1, 1e-5
    ny = H + 2
    cfg = cases.periodic_lbm_config(nx=5, ny=ny, nz=5, viscosity=nu, force=(Fx, 0.0, 0.0))
    solid = np.zeros((5, ny, 5), np.int32)
    solid[:, 0, :] = 1
    solid[:, -1, :] = 1
    s = po.OracleLBM(cfg, "f64")
    s.initialize(solid)
    for _ in range(6000):
        s.step()
    yj = np.arange(1, H + 1) - 0.5
    profile = (Fx / 3.0) / (2.0 * nu) * yj * (H - yj)
    got = s.rho[2, 1:-1, 2] * s.v[2, 1:-1, 2, 0] - Fx / 3.0
    # measured: 8e-8 of the peak (start-up not fully decayed after 6000 steps)
    assert np.abs(got - profile).max() < 1e-6 * profile.max(), (got, profile)
    # the Guo mass source -2/3 u.F varies with y and drives a cross flow of
    # ~5e-6 of the profile (the reference's, not a restatement error)
    assert np.abs(s.v[2, 1:-1, 2, 1:]).max() < 1e-4 * profile.max()
    assert np.all(s.v[:, 0] == 0.0) and np.all(s.v[:, -1] == 0.0)


def _apply_boundary(F, v, solid, left, right, rho_l, rho_r, vel_l, vel_r, rho0):
    nx, ny, nz = solid.shape
    po.lib().oref_lbm_apply_boundary_f64(
        nx, ny, nz, left, right, ctypes.c_double(rho_l), ctypes.c_double(rho_r), po._d3(vel_l),
        po._d3(vel_r), ctypes.c_double(rho0), po._p(solid), po._p(v), po._p(F))


def test_velocity_face_equals_its_net_form():
    """The in-place sequential update F_q <- feq(opp q) - F_opp q + feq(q) of
    fsi/lbm3d.py:374-382 is, over q = 0..18 with opp(2m-1) = 2m: F_0 <-
    2 feq_0 - F_0; odd q: F_q <- feq_q + feq_{q+1} - F_{q+1}; even q >= 2
    unchanged. It uses rho0, not the face's rho; solid face cells and interior
    cells are untouched."""
    rng = np.random.default_rng(8)
    n = (6, 5, 5)
    solid = (rng.uniform(size=n) < 0.25).astype(np.int32)
    F = rng.uniform(0.01, 0.1, (*n, 19))
    v = rng.uniform(-0.05, 0.05, (*n, 3))
    ul, ur, rho0 = np.array([0.04, 0.01, -0.02]), np.array([-0.03, 0.0, 0.02]), 1.05
    out = F.copy()
    _apply_boundary(out, v.copy(), solid, 2, 2, 7.0, 9.0, ul, ur, rho0)
    for x, ub in ((0, ul), (n[0] - 1, ur)):
        fe = feq(rho0, ub)
        net = F[x].copy()
        net[..., 0] = 2 * fe[0] - F[x][..., 0]
        for q in range(1, 19, 2):
            net[..., q] = fe[q] + fe[q + 1] - F[x][..., q + 1]
        fl = solid[x] == 0
        assert np.abs(out[x][fl] - net[fl]).max() < 1e-15
        assert np.array_equal(out[x][~fl], F[x][~fl])
    assert np.array_equal(out[1:-1], F[1:-1])
    # equilibrium at the face velocity is a fixed point
    eq = np.broadcast_to(feq(rho0, ul), (*n, 19)).copy()
    eq2 = eq.copy()
    _apply_boundary(eq2, v.copy(), np.zeros(n, np.int32), 2, 0, 1.0, 1.0, ul, ur, rho0)
    assert np.abs(eq2 - eq).max() < 1e-16


def test_pressure_face_reads_the_solid_neighbours_velocity():
    """fsi/lbm3d.py:365-372: a fluid face cell becomes feq(rho_face, u) with u
    the previous velocity of its inner neighbour if that neighbour is solid,
    else its own."""
    rng = np.random.default_rng(9)
    n = (6, 5, 5)
    solid = np.zeros(n, np.int32)
    solid[1] = (rng.uniform(size=n[1:]) < 0.5)
    solid[-2] = (rng.uniform(size=n[1:]) < 0.5)
    solid[0, 0, 0] = solid[-1, 4, 4] = 1
    F = rng.uniform(0.01, 0.1, (*n, 19))
    v = rng.uniform(-0.05, 0.05, (*n, 3))
    out = F.copy()
    _apply_boundary(out, v, solid, 1, 1, 1.02, 0.97, (0, 0, 0), (0, 0, 0), 1.0)
    for x, inner, rho_b in ((0, 1, 1.02), (n[0] - 1, n[0] - 2, 0.97)):
        for j in range(n[1]):
            for k in range(n[2]):
                if solid[x, j, k]:
                    assert np.array_equal(out[x, j, k], F[x, j, k])
                    continue
                src = v[inner, j, k] if solid[inner, j, k] else v[x, j, k]
                assert np.abs(out[x, j, k] - feq(rho_b, src)).max() < 1e-16
    assert solid[1].any() and not solid[1].all()


# ------------------------------------------------------------------- MPM --


def _one_particle(cfg, x, v=(0.0, 0.0, 0.0), F=np.eye(3), C=np.zeros((3, 3)), mass=0.8, vol=0.3):
    s = po.OracleMPM(cfg, "f64")
    s.initialize_particles_from_numpy(np.array([x], np.float32), np.array([v], np.float32),
                                      particle_mass=mass, particle_volume=vol)
    s.x[0] = x  # exact float64 position
    s.F[0] = F
    s.C[0] = C
    return s


def _p2g(s, dt):
    a = s.args(dt)
    s.grid_v[...] = 0
    s.grid_m[...] = 0
    s._call("mpm_p2g", ctypes.byref(a), po._p(s.x), po._p(s.v), po._p(s.C), po._p(s.F), po._p(s.mass),
            po._p(s.volume), po._p(s.active), po._p(s.particle_force), po._p(s.grid_v), po._p(s.grid_m))


def _node_offsets(s, x):
    """Node positions minus the particle position, for every node (dx units)."""
    idx = np.stack(np.meshgrid(np.arange(s.nx), np.arange(s.ny), np.arange(s.nz), indexing="ij"), -1)
    return idx * s.dx - np.asarray(x)


def test_p2g_weights_have_the_quadratic_bspline_moments():
    """Mass scattered by one particle: total m, centre at x_p, second moment
    m dx^2/4 I (the quadratic B-spline's), at random positions and dx."""
    rng = np.random.default_rng(2)
    for dx in (1.0, 0.5):
        cfg = cases.small_mpm_config(dx=dx, youngs_modulus=1.0)
        for _ in range(10):
            x = rng.uniform(4.0, 11.0, 3) * dx
            s = _one_particle(cfg, x)
            _p2g(s, 1e-3)
            d = _node_offsets(s, x)
            m, mp = s.grid_m, float(s.mass[0])  # the mass passes through float32
            assert m.sum() == pytest.approx(mp, rel=1e-14)
            assert np.abs(np.einsum("ijk,ijka->a", m, d)).max() < 1e-14
            second = np.einsum("ijk,ijka,ijkb->ab", m, d, d)
            assert np.abs(second - mp * dx * dx / 4 * np.eye(3)).max() < 1e-13


@pytest.mark.parametrize("dx", [1.0, 0.5])
def test_p2g_momentum_carries_the_fixed_corotated_stress(dx):
    """One P2G of a particle with C, F given: F is first updated to
    (I + dt C) F (fsi/mpm3d.py:243); the momentum scattered is
    w_i (m v + A (x_i - x_p) + dt f) with A = -dt V 4/dx^2 P + m C and
    P = 2 mu (F - R) F^T + lam J (J - 1) I, R the polar factor of F. The
    quadratic B-spline's moments give sum_i p_i = m v + dt f and
    sum_i p_i (x_i - x_p)^T = A dx^2/4, so A is recovered from the grid and
    compared with P from scipy's polar decomposition."""
    E_mod, nu_p = 1000.0, 0.3
    mu = E_mod / (2 * (1 + nu_p))
    lam = E_mod * nu_p / ((1 + nu_p) * (1 - 2 * nu_p))
    cfg = cases.small_mpm_config(dx=dx, youngs_modulus=E_mod, poisson_ratio=nu_p)
    rng = np.random.default_rng(6)
    dt = 2e-3
    for _ in range(6):
        Q1 = scipy.linalg.expm(np.cross(np.eye(3), rng.normal(0, 0.5, 3)))
        Q2 = scipy.linalg.expm(np.cross(np.eye(3), rng.normal(0, 0.5, 3)))
        F0 = Q1 @ np.diag(rng.uniform(0.8, 1.25, 3)) @ Q2
        C = rng.normal(0, 0.5, (3, 3))
        v = rng.normal(0, 0.1, 3)
        x = rng.uniform(5.0, 10.0, 3) * dx
        s = _one_particle(cfg, x, v=v, F=F0, C=C, mass=0.8, vol=0.3)
        s.v[0] = v
        f_p = rng.normal(0, 1e-2, 3)
        s.particle_force[0] = f_p
        _p2g(s, dt)
        mass, vol = float(s.mass[0]), float(s.volume[0])  # stored through float32
        Fn = (np.eye(3) + dt * C) @ F0
        assert np.abs(s.F[0] - Fn).max() < 1e-14
        R, _ = scipy.linalg.polar(Fn)
        J = np.linalg.det(Fn)
        P = 2 * mu * (Fn - R) @ Fn.T + lam * J * (J - 1) * np.eye(3)
        A = -dt * vol * 4 / dx ** 2 * P + mass * C
        d = _node_offsets(s, x)
        p = s.grid_v
        assert np.abs(p.sum((0, 1, 2)) - (mass * v + dt * f_p)).max() < 1e-13
        A_grid = np.einsum("ijka,ijkb->ab", p, d) * 4 / dx ** 2
        assert np.abs(A_grid - A).max() < 1e-11 * max(1.0, np.abs(A).max()), (A_grid, A)


def test_apic_transfers_an_affine_velocity_field_exactly():
    """Particles with v_p = A (x_p - x0) + b and C_p = A (no stress, no
    gravity, away from the walls): P2G gives every node v_i = A (x_i - x0) +
    b, G2P returns v_p and C_p = 4/dx sum w v_i (offset - fx)^T = A
    unchanged, and x moves by dt v."""
    cfg = cases.small_mpm_config(youngs_modulus=1e-30, dx=0.5, boundary_width=2, nx=24, ny=24, nz=24)
    rng = np.random.default_rng(12)
    pos = rng.uniform(4.0, 8.0, (300, 3))
    A = rng.normal(0, 0.2, (3, 3))
    b = np.array([0.05, -0.02, 0.03])
    x0 = np.array([6.0, 6.0, 6.0])
    s = po.OracleMPM(cfg, "f64")
    s.initialize_particles_from_numpy(pos.astype(np.float32), particle_mass=0.1, particle_volume=0.02)
    n = len(pos)
    s.x[:n] = pos
    v0 = (pos - x0) @ A.T + b
    s.v[:n] = v0
    s.C[:n] = A
    dt = 1e-3
    s.substep(dt)
    assert np.abs(s.v[:n] - v0).max() < 1e-13
    assert np.abs(s.C[:n] - A).max() < 1e-12
    assert np.abs(s.x[:n] - (pos + dt * v0)).max() < 1e-15


@pytest.mark.parametrize("damping", [0.0, 0.5])
def test_grid_walls_turn_outward_velocities_in_the_band(damping):
    """fsi/mpm3d.py:269-288 on random grid momenta/masses: v = p/m + dt g on
    nodes with mass; within boundary_width of a face, a velocity component
    pointing out of the domain becomes -damping times itself (-0.0 at zero
    damping); all other components, and massless nodes, are unchanged."""
    cfg = cases.small_mpm_config(nx=9, ny=10, nz=11, boundary_width=3, boundary_damping=damping,
                                 gravity=(0.01, -0.02, 0.005))
    s = po.OracleMPM(cfg, "f64")
    rng = np.random.default_rng(1)
    p = rng.normal(0, 1, (9, 10, 11, 3))
    m = rng.uniform(0.1, 2.0, (9, 10, 11))
    m[rng.uniform(size=m.shape) < 0.1] = 0.0
    s.grid_v[...] = p
    s.grid_m[...] = m
    dt = 0.25
    a = s.args(dt)
    s._call("mpm_grid_op", ctypes.byref(a), po._p(s.grid_v), po._p(s.grid_m))
    g = np.array([0.01, -0.02, 0.005])
    vel = p / np.where(m > 0, m, 1.0)[..., None] + dt * g
    dims = (9, 10, 11)
    flips = 0
    for ax in range(3):
        idx = np.arange(dims[ax]).reshape([-1 if a_ == ax else 1 for a_ in range(3)])
        out_low = (idx < 3) & (vel[..., ax] < 0)
        out_high = (idx >= dims[ax] - 3) & (vel[..., ax] > 0)
        flip = (out_low | out_high) & (m > 0)
        keep = ~flip & (m > 0)
        got = s.grid_v[..., ax]
        assert np.array_equal(got[flip], -damping * vel[..., ax][flip])
        assert np.abs(got[keep] - vel[..., ax][keep]).max() < 1e-15
        assert np.array_equal(got[m == 0], p[..., ax][m == 0])
        if damping == 0.0:
            # -0.0 * v: +0.0 where v < 0 (low faces), -0.0 where v > 0 (high)
            assert not np.any(np.signbit(got[flip & out_low]))
            assert np.all(np.signbit(got[flip & out_high]))
            assert np.all(got[flip] == 0.0)
        else:
            assert np.all(np.sign(got[flip]) == -np.sign(vel[..., ax][flip]))
        flips += int(flip.sum())
    assert flips > 100


def test_positions_clamp_to_the_domain():
    """fsi/mpm3d.py:321-328: after x += dt v each axis is clamped to
    [0.5 dx, (n - 1.5) dx]; particles pushed out stop exactly at the bound,
    the others move freely."""
    cfg = cases.small_mpm_config(nx=12, ny=14, nz=16, dx=0.5, boundary_width=1, boundary_damping=1.0)
    rng = np.random.default_rng(3)
    n = 400
    pos = rng.uniform(0.3, 5.5, (n, 3))
    vel = rng.normal(0, 4.0, (n, 3))
    s = po.OracleMPM(cfg, "f64")
    s.initialize_particles_from_numpy(pos.astype(np.float32), vel.astype(np.float32),
                                      particle_mass=0.1, particle_volume=0.01)
    s.x[:n] = np.clip(pos, 0.25, 5.25)
    x0 = s.x[:n].copy()
    dt = 0.1
    s.substep(dt)
    lo, hi = 0.25, np.array([(12 - 1.5) * 0.5, (14 - 1.5) * 0.5, (16 - 1.5) * 0.5])
    free = x0 + dt * s.v[:n]
    assert np.array_equal(s.x[:n], np.minimum(np.maximum(free, lo), hi))
    assert np.any(free < lo) and np.any(free > hi)
    assert np.all(s.x[:n] >= lo) and np.all(s.x[:n] <= hi)


def _oracle_coupling(refine, contact):
    from parity import REFINED_MASS, REFINED_VOLUME, refined_scene

    cfg, solid, pos, vel = refined_scene(refine=refine, contact=contact, ib=True)
    s = po.OracleSimulation(cfg, "f64")
    s.initialize_lbm(solid)
    s.initialize_mpm_from_numpy(pos, vel, particle_mass=REFINED_MASS, particle_volume=REFINED_VOLUME)
    s.coupler.compute_coupling_forces(0.25)
    c = s.coupler
    return {
        "cforce": c.coupling_force_numpy(), "phi": c.solid_volume_fraction_numpy(),
        "ibf": c.immersed_boundary_force_numpy(), "pforce": s.mpm.particle_forces_numpy(),
        "wvalid": c.particle_valid_weights_numpy(), "cmask": c.particle_coupling_mask_numpy(),
        "contact": c.particle_contact_mask_numpy(), "counters": np.array(c.counters).copy(),
    }


def test_refined_grid_couples_on_the_lattice():
    """The coupling (fsi/coupling.py:463-496) reads the lattice and the
    particles only: with the MPM grid twice finer than the lattice, the
    same particles give the same coupling force, solid fraction, particle
    forces, masks and counters as on a co-located MPM grid, and an MPM-grid
    stencil (node spacing half a lattice cell) would not."""
    a = _oracle_coupling(2, True)
    b = _oracle_coupling(1, True)
    assert int(a["counters"][3]) > 0 and float(np.abs(a["cforce"]).max()) > 0
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    # the fluid reaction balances the particle forces, weighted by the
    # substep's share of the lattice step (dt_ratio 0.25; all cells are fluid
    # where the cube sits)
    tot_p = a["pforce"].sum(axis=0)
    tot_f = a["cforce"].reshape(-1, 3).sum(axis=0)
    assert np.abs(0.25 * tot_p + tot_f).max() < 1e-12 * np.abs(tot_p).max()
