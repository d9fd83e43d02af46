"""MPM grid finer than the lattice (CouplingConfig.mpm_refinement = 2: the
MPM grid spacing is half a lattice cell, as in the benchmark configurations
that give dx_mpm = dx_lbm / 2). The coupling stencil then runs on the
lattice and the P2G/G2P stencil on the MPM grid, in the same fused kernels.

Gates as in test_gpu_mpm_coupling_parity.py: the FP64 build matches the
FP64 oracle to 1e-9 relative per field (float atomics reorder the sums);
FP32 stays within twice the reference-order FP32 error band against FP64
(+1e-6); masks and counters are exact. Each schedule the launchers pick is
run: sorted particles with the plane split (three wavefronts per particle,
the default at this size; FP64 runs one wavefront per particle), the
one-batch and the per-row gathers, shuffled unsorted particles (node boxes
too large for the LDS tile: global atomics), and the two-launch substeps. The coupling
alone (compute_coupling_forces) must equal that of the same particles on a
co-located MPM grid: it reads only the lattice and the particles.
"""

import numpy as np
import pytest

from oracle import pyoracle as po
from parity import REFINED_MASS, REFINED_VOLUME, close, refined_scene, rel_inf, sim_fields

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("gpu_available")]

SCENES = {
    "penalty": dict(steps=6),
    "ib_contact_wall": dict(steps=4, ib=True, contact=True, wall=True),
}
FIELDS = ("dev", "rho", "u", "x", "v", "F", "pforce", "cforce", "phi", "lbm_force", "wvalid")
COUNTS = ("unsupported_particle_count", "partial_support_particle_count", "clipped_particle_count",
          "ib_active_cell_count", "contact_candidate_count", "contact_damped_particle_count")


def _prepare(sim, schedule):
    if schedule in ("wide", "per_row"):
        sim.mpm.split_max_chunks = 0
    if schedule == "per_row":
        sim.mpm.wide_max_workgroups = 0
    elif schedule == "unsorted":
        sim.coupler.sort_interval = 0
    elif schedule == "two_launch":
        make = sim.mpm._make_buffers

        def patched():
            b = make()
            b.grid_alt2 = None
            return b

        sim.mpm._make_buffers = patched
        sim.mpm._buffers = patched()


PER_PARTICLE = ("x", "v", "F", "pforce", "cmask", "wvalid")


def _run(make, storage, steps, schedule=None, **kw):
    cfg, solid, pos, vel = refined_scene(storage=storage, **kw)
    perm = None
    if schedule == "unsorted":
        # shuffled and never sorted: every workgroup's node box spans the
        # cube, beyond the LDS tile
        perm = np.random.default_rng(11).permutation(len(pos))
        pos, vel = pos[perm], vel[perm]
    sim = make(cfg)
    sim.initialize_lbm(solid)
    sim.initialize_mpm_from_numpy(pos, vel, particle_mass=REFINED_MASS, particle_volume=REFINED_VOLUME)
    if schedule is not None:
        _prepare(sim, schedule)
    for _ in range(steps):
        sim.step()
    out = sim_fields(sim)
    out["diag"] = sim.coupler.coupling_diagnostics()
    if perm is not None:
        for key in PER_PARTICLE:
            back = np.empty_like(out[key])
            back[perm] = out[key]
            out[key] = back
    return out


@pytest.fixture(scope="module")
def oracle_runs():
    return {name: {p: _run(lambda c, p=p: po.OracleSimulation(c, p), None, **kw) for p in ("f32", "f64")}
            for name, kw in SCENES.items()}


@pytest.mark.parametrize("schedule", ["split", "wide", "per_row", "unsorted", "two_launch"])
@pytest.mark.parametrize("name", list(SCENES))
def test_refined_fp64_matches_fp64_oracle(name, schedule, oracle_runs):
    import fsi

    got = _run(fsi.FSISimulation, "fp64", schedule=schedule, **SCENES[name])
    ref = oracle_runs[name]["f64"]
    for key in FIELDS:
        ok, err, diff = close(got[key], ref[key], 1e-9, 1e-13)
        print(f"{name}/{schedule} fp64 {key}: rel {err:.3e} abs {diff:.3e}")
        assert ok, (name, schedule, key, err, diff)
    assert np.array_equal(got["cmask"], ref["cmask"])
    for k in COUNTS:
        assert got["diag"][k] == ref["diag"][k], (k, got["diag"][k], ref["diag"][k])
    if name == "ib_contact_wall":
        assert ref["diag"]["partial_support_particle_count"] > 0
        assert ref["diag"]["contact_damped_particle_count"] > 0


@pytest.mark.parametrize("schedule", ["split", "wide", "per_row"])
@pytest.mark.parametrize("name", list(SCENES))
def test_refined_fp32_within_reference_band(name, schedule, oracle_runs):
    import fsi

    got = _run(fsi.FSISimulation, "fp32", schedule=schedule, **SCENES[name])
    ref64, ref32 = oracle_runs[name]["f64"], oracle_runs[name]["f32"]
    for key in FIELDS:
        e_gpu = rel_inf(got[key], ref64[key])
        e_ref = rel_inf(ref32[key], ref64[key])
        print(f"{name}/{schedule} fp32 {key}: gpu-vs-f64 {e_gpu:.3e} ref32-vs-f64 {e_ref:.3e}")
        assert e_gpu <= 2.0 * e_ref + 1e-6, (name, schedule, key, e_gpu, e_ref)
    assert np.array_equal(got["cmask"], ref32["cmask"])


def _coupling_only(make, storage, refine):
    cfg, solid, pos, vel = refined_scene(storage=storage, refine=refine, ib=True, contact=True, wall=True)
    sim = make(cfg)
    sim.initialize_lbm(solid)
    sim.initialize_mpm_from_numpy(pos, vel, particle_mass=REFINED_MASS, particle_volume=REFINED_VOLUME)
    sim.coupler.compute_coupling_forces(0.25)
    c = sim.coupler
    return {
        "cforce": c.coupling_force_numpy(), "phi": c.solid_volume_fraction_numpy(),
        "pforce": sim.mpm.particle_forces_numpy(), "wvalid": c.particle_valid_weights_numpy(),
        "cmask": c.particle_coupling_mask_numpy(), "contact": c.particle_contact_mask_numpy(),
        "diag": c.coupling_diagnostics(),
    }


@pytest.mark.parametrize("storage", ["fp64", "fp32"])
def test_refined_coupling_equals_the_co_located_one(storage):
    import fsi

    fine = _coupling_only(fsi.FSISimulation, storage, 2)
    flat = _coupling_only(fsi.FSISimulation, storage, 1)
    ref = _coupling_only(lambda c: po.OracleSimulation(c, "f64"), None, 2)
    # FP32: the two grids group the particles into different workgroups, so
    # the float atomics that flush each workgroup's tile add in another
    # order (a few ulp of the largest value: 1.0e-6 measured)
    tol = 1e-12 if storage == "fp64" else 4e-6
    for key in ("cforce", "phi", "pforce", "wvalid"):
        e_flat = rel_inf(fine[key], flat[key])
        e_ref = rel_inf(fine[key], ref[key])
        print(f"{storage} {key}: refined vs co-located {e_flat:.3e}, vs oracle {e_ref:.3e}")
        assert e_flat <= tol, (key, e_flat)
        assert e_ref <= (1e-9 if storage == "fp64" else 1e-5), (key, e_ref)
    for key in ("cmask", "contact"):
        assert np.array_equal(fine[key], flat[key]) and np.array_equal(fine[key], ref[key]), key
    for k in ("unsupported_particle_count", "partial_support_particle_count", "clipped_particle_count",
              "contact_candidate_count", "contact_damped_particle_count"):
        assert fine["diag"][k] == flat["diag"][k] == ref["diag"][k], (k, fine["diag"][k], ref["diag"][k])
    assert ref["diag"]["partial_support_particle_count"] > 0


def test_refined_grid_rejects_the_closure():
    import fsi

    cfg, _, _, _ = refined_scene(storage="fp32")
    sim = fsi.FSISimulation(cfg)
    with pytest.raises(NotImplementedError):
        sim.coupler.set_closure(object())


def test_refined_by_three_fp64_matches_fp64_oracle():
    """Three MPM cells per lattice cell and axis (MPM dx = 1/3): the
    refinement is not tied to powers of two."""
    import fsi

    kw = dict(steps=4, refine=3, substeps=24, ib=True, contact=True, wall=True)
    ref = _run(lambda c: po.OracleSimulation(c, "f64"), None, **kw)
    got = _run(fsi.FSISimulation, "fp64", **kw)
    for key in FIELDS:
        ok, err, diff = close(got[key], ref[key], 1e-9, 1e-13)
        print(f"refine 3 fp64 {key}: rel {err:.3e} abs {diff:.3e}")
        assert ok, (key, err, diff)
    assert np.array_equal(got["cmask"], ref["cmask"])
    for k in COUNTS:
        assert got["diag"][k] == ref["diag"][k], (k, got["diag"][k], ref["diag"][k])
