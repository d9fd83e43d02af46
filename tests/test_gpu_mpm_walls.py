"""MPM grid walls, position clamp and the large-set kernel variants against the oracle.

Scenes in which particles enter the boundary_width band of the grid (so the
sticky/separating wall of fsi/mpm3d.py:269-288 turns their grid velocities)
and reach the position clamp of fsi/mpm3d.py:321-328 on all six faces, at
boundary_damping 0 and 0.5: an elastic cube thrown at a corner plus six lone
particles thrown at the six faces.

The particle passes have two gather variants: launches of at most
``wide_max_workgroups`` workgroups gather all 27 stencil nodes in one batch,
larger ones one x plane at a time, and the FP32 fused G2P+P2G kernel of
large particle sets is held to 128 VGPRs with its G2P stage and P2G tile in
one LDS buffer. Scenes small enough for the oracle are forced onto the
large-set variant with ``wide_max_workgroups = 0`` (and the FP32 plane split
off), so the kernels of the big configurations are compared with the oracle
itself, not with another build of the same code.

Gates as in test_gpu_mpm_coupling_parity.py: FP64 build vs FP64 oracle to
1e-9 relative per field; FP32 within twice the reference-order FP32 oracle's
error against FP64; clamp counts and coupling masks exact.
"""

import dataclasses

import numpy as np
import pytest

from cases import small_mpm_config
from oracle import pyoracle as po
from parity import close, coupled_scene, jittered_cube, rel_inf, sim_fields

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("gpu_available")]

LONE_SPEED = 4.0


def _throw(n, lone_x):
    """A jittered cube near the (x-, y+, z-) corner moving into it, and six
    lone particles next to the six faces moving out through them."""
    cube = jittered_cube((1.0, n[1] - 6.5, 3.0), (4.0, n[1] - 3.5, 6.0), 0.5, seed=2)
    vc = np.tile(np.array([-1.5, 1.0, -0.5], np.float32), (len(cube), 1))
    cx, cy, cz = lone_x
    lone = np.array([[1.3, 5.0, cz], [n[0] - 1.8, 5.0, cz], [cx, 1.3, cz], [cx, n[1] - 1.8, cz],
                     [cx + 4.0, 10.0, 1.3], [cx - 6.0, 6.0, n[2] - 1.8]], np.float32)
    vl = LONE_SPEED * np.array([[-1, 0, 0], [1, 0, 0], [0, -1, 0], [0, 1, 0], [0, 0, -1], [0, 0, 1]],
                               np.float32)
    return np.concatenate([cube, lone]), np.concatenate([vc, vl])


def _mpm_scene(bw, damping):
    cfg = small_mpm_config(youngs_modulus=20.0, poisson_ratio=0.3, dt=0.05, boundary_width=bw,
                           boundary_damping=damping, gravity=(0.0, 0.0, -0.5))
    pos, vel = _throw((16, 16, 16), (10.0, 10.0, 12.0))
    return cfg, pos, vel


def _clamped(x, n, dx=1.0):
    """Particles at the lower / upper clamp per axis."""
    hi = (np.array(n, np.float64) - 1.5) * dx
    return (x == 0.5 * dx).sum(0), (x == hi).sum(0)


def _run_mpm(solver, pos, vel, steps, sort=False, wide=None):
    """Runs the substeps; "clamps" holds, per face, the largest number of
    particles found at the clamp after any substep."""
    solver.initialize_particles_from_numpy(pos, vel, particle_mass=0.125, particle_volume=0.125)
    if wide is not None:
        solver.wide_max_workgroups = wide
        solver.split_max_chunks = 0  # FP32: no plane split either
    if sort:
        solver.sort_particles()
    clamps = np.zeros(6, int)
    for _ in range(steps):
        solver.substep()
        clamps = np.maximum(clamps, np.concatenate(_clamped(solver.positions_numpy(), (16, 16, 16))))
    return {"x": solver.positions_numpy(), "v": solver.velocities_numpy(),
            "F": solver.deformation_gradients_numpy(), "clamps": clamps}


@pytest.fixture(scope="module")
def wall_oracle():
    out = {}
    for bw in (1, 3):
        for damping in (0.0, 0.5):
            cfg, pos, vel = _mpm_scene(bw, damping)
            out[bw, damping] = {p: _run_mpm(po.OracleMPM(cfg, p), pos, vel, 40) for p in ("f32", "f64")}
    return out


# variant: (sort first, wide_max_workgroups); unsorted: one workgroup whose box
# exceeds the LDS tile (global scatter and gathers); sorted: per-chunk tiles
VARIANTS = {"unsorted_wide": (False, None), "sorted_plane_batches": (True, 0)}


@pytest.mark.parametrize("variant", list(VARIANTS))
@pytest.mark.parametrize("damping", [0.0, 0.5])
@pytest.mark.parametrize("bw", [1, 3])
def test_mpm_walls_fp64_match_the_oracle(bw, damping, variant, wall_oracle):
    import fsi

    cfg, pos, vel = _mpm_scene(bw, damping)
    sort, wide = VARIANTS[variant]
    got = _run_mpm(fsi.MPMSolver3D(cfg, precision="fp64"), pos, vel, 40, sort, wide)
    ref = wall_oracle[bw, damping]["f64"]
    for key in ("x", "v", "F"):
        ok, err, diff = close(got[key], ref[key], 1e-9, 1e-13)
        print(f"walls bw={bw} d={damping} {variant} fp64 {key}: rel {err:.3e} abs {diff:.3e}")
        assert ok, (key, err, diff)
    assert np.array_equal(got["clamps"], ref["clamps"]), (got["clamps"], ref["clamps"])
    if bw == 1:  # the lone particles reach the clamp of every face
        assert np.all(got["clamps"] >= 1), got["clamps"]


@pytest.mark.parametrize("damping", [0.0, 0.5])
@pytest.mark.parametrize("bw", [1, 3])
def test_mpm_walls_fp32_within_reference_band(bw, damping, wall_oracle):
    import fsi

    cfg, pos, vel = _mpm_scene(bw, damping)
    got = _run_mpm(fsi.MPMSolver3D(cfg), pos, vel, 40, sort=True, wide=0)
    for key in ("x", "v", "F"):
        e_gpu = rel_inf(got[key], wall_oracle[bw, damping]["f64"][key])
        e_ref = rel_inf(wall_oracle[bw, damping]["f32"][key], wall_oracle[bw, damping]["f64"][key])
        print(f"walls bw={bw} d={damping} fp32 {key}: gpu-vs-f64 {e_gpu:.3e} ref32-vs-f64 {e_ref:.3e}")
        assert e_gpu <= 2.0 * e_ref + 2e-6, (key, e_gpu, e_ref)


# ---------------------------------------------------------------- coupled --


def _coupled_wall_scene(damping, storage=None):
    """The small coupled channel (24x16x16) with the thrown cube and lone
    particles; boundary_width 1 so the lone particles reach the clamp, MPM
    dt 0.05 (lbm_dt 0.4, 8 substeps) with a soft solid (E = 20). The penalty
    stiffness is 0.01: the particles move at 1.5-4 cells per unit time, and
    at the default 0.5 their reaction drives the fluid far past the lattice
    speed of sound (|u| ~ 300 after six steps); at 0.01 |u| stays below 0.04."""
    cfg, solid, _, _ = coupled_scene(storage=storage, n=(24, 16, 16), substeps=8, e_mod=20.0, gamma=0.01)
    mpm = dataclasses.replace(cfg.mpm, boundary_width=1, boundary_damping=damping)
    cfg = dataclasses.replace(cfg, lbm_dt=0.4, mpm=mpm)
    pos, vel = _throw((24, 16, 16), (14.0, 10.0, 11.0))
    return cfg, solid, pos, vel


def _run_coupled(make, scene, steps, tuning=None):
    cfg, solid, pos, vel = scene
    sim = make(cfg)
    if tuning:
        for k, v in tuning.items():
            setattr(sim.mpm, k, v)
    sim.initialize_lbm(solid)
    sim.initialize_mpm_from_numpy(pos, vel, particle_mass=0.15, particle_volume=0.125)
    clamps = np.zeros(6, int)
    for _ in range(steps):
        sim.step()
        n = (sim.mpm.nx, sim.mpm.ny, sim.mpm.nz)
        clamps = np.maximum(clamps, np.concatenate(_clamped(sim.mpm.positions_numpy(), n)))
    out = sim_fields(sim)
    out["diag"] = sim.coupler.coupling_diagnostics()
    out["clamps"] = clamps
    return out


# the large-set kernels: plane-batch gathers, no plane split
LARGE = {"wide_max_workgroups": 0, "split_max_chunks": 0}
FIELDS = ("dev", "rho", "u", "x", "v", "F", "pforce", "cforce", "phi", "lbm_force", "wvalid")
COUNTERS = ("unsupported_particle_count", "partial_support_particle_count", "clipped_particle_count",
            "ib_active_cell_count", "contact_candidate_count", "contact_damped_particle_count")


@pytest.fixture(scope="module")
def coupled_wall_oracle():
    return {d: {p: _run_coupled(lambda c, p=p: po.OracleSimulation(c, p), _coupled_wall_scene(d), 6)
                for p in ("f32", "f64")} for d in (0.0, 0.5)}


@pytest.mark.parametrize("kernels", ["default", "large_set"])
@pytest.mark.parametrize("damping", [0.0, 0.5])
def test_coupled_walls_fp64_match_the_oracle(damping, kernels, coupled_wall_oracle):
    import fsi

    got = _run_coupled(fsi.FSISimulation, _coupled_wall_scene(damping, "fp64"), 6,
                       LARGE if kernels == "large_set" else None)
    ref = coupled_wall_oracle[damping]["f64"]
    for key in FIELDS:
        ok, err, diff = close(got[key], ref[key], 1e-9, 1e-13)
        print(f"coupled walls d={damping} {kernels} fp64 {key}: rel {err:.3e} abs {diff:.3e}")
        assert ok, (key, err, diff)
    assert np.array_equal(got["cmask"], ref["cmask"])
    for k in COUNTERS:
        assert got["diag"][k] == ref["diag"][k], (k, got["diag"][k], ref["diag"][k])
    assert np.array_equal(got["clamps"], ref["clamps"]), (got["clamps"], ref["clamps"])
    assert np.all(got["clamps"] >= 1), got["clamps"]


@pytest.mark.parametrize("kernels", ["default", "large_set"])
@pytest.mark.parametrize("damping", [0.0, 0.5])
def test_coupled_walls_fp32_within_reference_band(damping, kernels, coupled_wall_oracle):
    import fsi

    got = _run_coupled(fsi.FSISimulation, _coupled_wall_scene(damping, "fp32"), 6,
                       LARGE if kernels == "large_set" else None)
    ref64, ref32 = coupled_wall_oracle[damping]["f64"], coupled_wall_oracle[damping]["f32"]
    for key in FIELDS:
        e_gpu = rel_inf(got[key], ref64[key])
        e_ref = rel_inf(ref32[key], ref64[key])
        print(f"coupled walls d={damping} {kernels} fp32 {key}: gpu-vs-f64 {e_gpu:.3e} ref32-vs-f64 {e_ref:.3e}")
        assert e_gpu <= 2.0 * e_ref + 1e-6, (key, e_gpu, e_ref)
    assert np.array_equal(got["cmask"], ref32["cmask"])


@pytest.fixture(scope="module")
def ib_contact_oracle():
    scene = coupled_scene(ib=True, contact=True)
    return {p: _run_coupled(lambda c, p=p: po.OracleSimulation(c, p), scene, 4) for p in ("f32", "f64")}


@pytest.mark.parametrize("storage", ["fp64", "fp32"])
def test_large_set_kernels_with_ib_and_contact_match_the_oracle(storage, ib_contact_oracle):
    """The coupled cube scene with immersed-boundary drag and contact on the
    large-set kernels (k_g2p_p2g with one x plane of gathers per batch; in
    FP32 the 128-VGPR variant with the shared LDS buffer)."""
    import fsi

    got = _run_coupled(fsi.FSISimulation, coupled_scene(storage=storage, ib=True, contact=True), 4, LARGE)
    ref64, ref32 = ib_contact_oracle["f64"], ib_contact_oracle["f32"]
    for key in FIELDS:
        if storage == "fp64":
            ok, err, diff = close(got[key], ref64[key], 1e-9, 1e-13)
            print(f"large-set ib/contact fp64 {key}: rel {err:.3e} abs {diff:.3e}")
            assert ok, (key, err, diff)
        else:
            e_gpu = rel_inf(got[key], ref64[key])
            e_ref = rel_inf(ref32[key], ref64[key])
            print(f"large-set ib/contact fp32 {key}: gpu-vs-f64 {e_gpu:.3e} ref32-vs-f64 {e_ref:.3e}")
            assert e_gpu <= 2.0 * e_ref + 1e-6, (key, e_gpu, e_ref)
    ref = ref64 if storage == "fp64" else ref32
    assert np.array_equal(got["cmask"], ref["cmask"])
    if storage == "fp64":
        for k in COUNTERS:
            assert got["diag"][k] == ref["diag"][k], (k, got["diag"][k], ref["diag"][k])
