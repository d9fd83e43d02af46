"""fsi_mpm_substeps_noalloc against fsi_mpm_substeps, through the C ABI.

The allocation-free entry point runs the same kernels with the node-box and
counter-row scratch in the library's static device array. Two coupled scenes
in a row (the second with fewer particles, so the static array holds the
first call's larger contents when it starts): every coupled step's substeps
go through one entry point or the other, and the FP64 fields must agree to
rounding of the float atomics' order (1e-10 relative), counters exactly. A
third scene runs the MPM grid at half the lattice spacing.
"""

import ctypes

import numpy as np
import pytest

from parity import REFINED_MASS, REFINED_VOLUME, close, coupled_scene, jittered_cube, refined_scene, sim_fields

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("gpu_available")]


def _two_cube_scene():
    cfg, solid, _, _ = coupled_scene(storage="fp64", n=(40, 28, 28), substeps=6, ib=True, contact=True)
    a = jittered_cube((4.0, 4.0, 4.0), (8.0, 8.0, 8.0), 0.5, 3)
    b = jittered_cube((25.0, 15.0, 15.0), (31.0, 21.0, 21.0), 0.5, 4)
    pos = np.concatenate([a, b])
    vel = np.zeros_like(pos)
    vel[: len(a), 0] = 0.01
    vel[len(a):, 1] = -0.01
    return cfg, solid, pos, vel


def _run(scene, entry, steps, mass=1.2 * 0.125, volume=0.125):
    import fsi
    from fsi import _native as nat

    cfg, solid, pos, vel = scene
    sim = fsi.FSISimulation(cfg)
    sim.initialize_lbm(solid)
    sim.initialize_mpm_from_numpy(pos, vel, particle_mass=mass, particle_volume=volume)
    real_call = nat.call
    seen = []

    def call(name, *args):
        if name == "fsi_mpm_substeps":
            name = entry
            seen.append(name)
        return real_call(name, *args)

    nat.call = call
    try:
        for _ in range(steps):
            sim.step()
    finally:
        nat.call = real_call
    assert seen == [entry] * steps
    out = sim_fields(sim)
    out["diag"] = sim.coupler.coupling_diagnostics()
    return out


FIELDS = ("dev", "rho", "u", "x", "v", "F", "pforce", "cforce", "phi", "lbm_force", "wvalid")


def test_noalloc_substeps_match_the_allocating_entry_point():
    plain = {}
    scenes = [("two cubes", _two_cube_scene(), plain),
              ("one cube", coupled_scene(storage="fp64", ib=True, contact=True), plain),
              # the MPM grid at half the lattice spacing (the refined kernels)
              ("refined", refined_scene(storage="fp64", ib=True, contact=True, wall=True),
               {"mass": REFINED_MASS, "volume": REFINED_VOLUME})]
    assert len(scenes[1][1][2]) < len(scenes[0][1][2])
    for label, scene, kw in scenes:
        ref = _run(scene, "fsi_mpm_substeps", 3, **kw)
        got = _run(scene, "fsi_mpm_substeps_noalloc", 3, **kw)
        for key in FIELDS:
            ok, err, diff = close(got[key], ref[key], 1e-10, 1e-14)
            print(f"noalloc {label} {key}: rel {err:.3e} abs {diff:.3e}", flush=True)
            assert ok, (label, key, err, diff)
        assert np.array_equal(got["cmask"], ref["cmask"])
        assert got["diag"] == pytest.approx(ref["diag"], rel=1e-10, abs=1e-14)
        for k in ("unsupported_particle_count", "partial_support_particle_count", "clipped_particle_count",
                  "ib_active_cell_count", "contact_candidate_count", "contact_damped_particle_count"):
            assert got["diag"][k] == ref["diag"][k], (label, k)


def test_noalloc_substeps_reject_work_beyond_the_static_scratch():
    """More particle workgroups than the static array holds: a clear error,
    nothing launched."""
    import fsi
    from fsi import _native as nat

    cfg, solid, pos, vel = coupled_scene(storage="fp64")
    sim = fsi.FSISimulation(cfg)
    sim.initialize_lbm(solid)
    sim.initialize_mpm_from_numpy(pos, vel, particle_mass=0.15, particle_volume=0.125)
    mpm = sim.mpm
    p = mpm.params(mpm.dt_default)
    p.num_particles = mpm.max_particles
    p.num_chunks = 70000  # work list length beyond 65536 workgroups
    b = mpm._make_buffers()
    b.chunks = ctypes.c_void_p(mpm.x.data_ptr())  # any non-null pointer: never read
    lib = nat.load_library()
    rc = lib.fsi_mpm_substeps_noalloc(nat.ref(p), nat.ref(b), None, None, 4, 0, 1, None)
    assert rc == -1
    msg = lib.fsi_last_error().decode()
    assert "exceed the static scratch" in msg, msg
