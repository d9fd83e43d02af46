"""Properties of the full-size coupled suspension (256x128x128 lattice, 16
elastic spheres of log-uniform E; bench.coupled_scene_large), on the
large-set kernel variants: on a co-located MPM grid (448,861 particles, 256
substeps per lattice step) and on the 512x256x256 MPM grid at
dx_mpm = dx_lbm / 2 (3,591,333 particles, 512 substeps).

No oracle runs at this size; the checks are the ones the physics fixes:

* momentum balance of the penalty coupling (fsi/coupling.py:332-406): the
  reaction spread on the lattice equals minus the particle force on every
  substep, so over a step the fluid's coupling force density summed over
  fluid cells (times the cell volume, 1) cancels the particles' step-averaged
  coupling force (their force times dt_ratio, summed over substeps);
* the flow stays bounded and finite (inflow 0.04) and the spheres move
  less than a cell per step;
* lattice mass: the FP64 build's total mass (an FP64 device sum) changes by
  the open faces' net flux only, well under 1e-3 of the total over 8 steps,
  and the FP32 and mixed-precision (I16 populations) builds agree with it.
"""

import numpy as np
import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("gpu_available")]

STEPS = 8


PARTICLES = {1: 448861, 2: 3591333}


def _run(storage, refine):
    import bench

    sim, cfg, pos = bench.large_sim(storage, refine)
    sim.coupler.record_step_force()
    m0 = sim.diagnostics()["lbm_total_mass"]
    x0 = sim.mpm.positions_numpy().astype(np.float64)
    for _ in range(STEPS):
        sim.step()
    d = sim.diagnostics()
    out = {
        "m0": m0, "m": d["lbm_total_mass"],
        "u": sim.lbm.velocity_numpy().astype(np.float64),
        "v": sim.mpm.velocities_numpy().astype(np.float64),
        "x": sim.mpm.positions_numpy().astype(np.float64), "x0": x0,
        "step_force": sim.coupler.step_force_numpy().astype(np.float64).sum(0),
        "fluid_force": sim.coupler.coupling_force_numpy().astype(np.float64)[sim.lbm.solid_numpy() == 0].sum(0),
        "particles": len(pos), "unsupported": d["coupling_unsupported_particle_count"],
    }
    del sim
    import torch

    torch.cuda.empty_cache()
    return out


_cache = {}


@pytest.fixture(scope="module", params=[1, 2], ids=["colocated", "dx_half"])
def runs(request):
    refine = request.param
    if refine not in _cache:
        _cache.clear()  # one scene's runs held at a time
        _cache[refine] = {s: _run(s, refine) for s in ("fp64", "fp32", "i16")}
        _cache[refine]["refine"] = refine
    return _cache[refine]


@pytest.mark.parametrize("storage", ["fp64", "fp32", "i16"])
def test_coupling_momentum_balance(runs, storage):
    r = runs[storage]
    fp, ff = r["step_force"], r["fluid_force"]
    scale = np.abs(fp).max()
    print(f"{storage}: particles {fp}, fluid {ff}")
    assert scale > 0.0
    tol = 1e-10 if storage == "fp64" else 1e-4
    assert np.abs(fp + ff).max() <= tol * scale, (fp, ff)


@pytest.mark.parametrize("storage", ["fp64", "fp32", "i16"])
def test_flow_and_bodies_stay_bounded(runs, storage):
    r = runs[storage]
    assert r["particles"] == PARTICLES[runs["refine"]] and r["unsupported"] == 0
    for k in ("u", "v", "x"):
        assert np.all(np.isfinite(r[k])), k
    umax = np.linalg.norm(r["u"], axis=-1).max()
    vmax = np.linalg.norm(r["v"], axis=-1).max()
    disp = np.abs(r["x"] - r["x0"]).max()
    print(f"{storage}: |u| max {umax:.4f}, |v| max {vmax:.4f}, displacement {disp:.4f} cells")
    assert 0.03 < umax < 0.2
    assert vmax < 0.2
    assert disp < STEPS * 1.0


def test_lattice_mass(runs):
    ref = runs["fp64"]
    drift = abs(ref["m"] - ref["m0"]) / ref["m0"]
    print(f"fp64 mass {ref['m0']:.10e} -> {ref['m']:.10e}, relative change {drift:.2e}")
    assert drift < 1e-3
    for storage in ("fp32", "i16"):
        e = abs(runs[storage]["m"] - ref["m"]) / ref["m"]
        print(f"{storage}: mass vs fp64 {e:.2e}")
        assert e < 1e-5, (storage, e)
    # the mixed-precision flow field stays close to the FP64 one
    for storage in ("fp32", "i16"):
        e = np.abs(runs[storage]["u"] - ref["u"]).max() / np.abs(ref["u"]).max()
        print(f"{storage}: u vs fp64 {e:.2e}")
        assert e < 1e-2, (storage, e)
