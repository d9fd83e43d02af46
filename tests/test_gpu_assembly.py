"""The coupled step's force assembly fused into the lattice collision
(fsi_lbm_collide_assembled) against the two-pass sequence it replaces: the
assembly pass (fsi_force_assemble: fsi/coupling.py:162-170, 408-461) and
then the collide phase with the assembled force field.

The fused kernel evaluates the same expressions per cell. Two runs of the
same coupled scene are not bit-identical (the P2G sums use atomics, so their
order varies), so each field of the two-pass run is compared with the fused
run against the measured difference of two fused runs (4x that floor, or
1e-12 / 1e-6 relative in FP64 / FP32 when the runs happen to agree), with
IB drag (and its clipping) and the solid-fraction clamp. I16 storage: its
adaptive 16-bit scale turns last-bit differences into whole quanta (two
fused runs already differ by several), so there the check is what the
fusion could break: the first writes run a measuring pass that must count
no IB cell (the counters equal the FP32 run's), and the fields agree with
the two-pass run to the storage's accuracy.
"""

import numpy as np
import pytest

from fsi import _native as nat
from parity import coupled_scene, sim_fields

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("gpu_available")]


def _two_pass(self, assembly, cforce, ib_force, ib_total, counters):
    # the sequence LBMMpmCoupler.step ran before the fusion
    self._before_force_update()
    nat.call("fsi_force_assemble", nat.ref(self._params), nat.ref(self._buffers), nat.ref(assembly),
             nat.ptr(cforce), nat.ptr(ib_force), nat.ptr(ib_total), nat.ptr(counters), 1)
    self._force_uniform = None
    self.step()


def _run(storage, fused, ib_max_force=None, steps=3):
    import fsi

    import dataclasses

    cfg, solid, pos, vel = coupled_scene(storage=storage, substeps=4, ib=True)
    cfg = dataclasses.replace(cfg, coupling=dataclasses.replace(cfg.coupling, immersed_boundary_max_force=ib_max_force))
    sim = fsi.FSISimulation(cfg)
    if not fused:
        sim.lbm.collide_assembled = _two_pass.__get__(sim.lbm)
    sim.initialize_lbm(solid)
    sim.initialize_mpm_from_numpy(pos, vel, particle_mass=0.15, particle_volume=0.125)
    for _ in range(steps):
        sim.step()
    out = sim_fields(sim)
    out["ibf"] = sim.coupler.immersed_boundary_force_numpy()
    out["ib_total"] = sim.coupler.ib_total.cpu().numpy().astype(np.float64)
    out["counters"] = sim.coupler.counters.cpu().numpy().copy()
    out["post"] = sim.lbm.populations.cpu().numpy().copy()
    return out


@pytest.mark.parametrize("storage,ib_max_force", [("fp64", None), ("fp32", None), ("fp32", 2.0e-4)])
def test_fused_assembly_equals_the_two_passes(storage, ib_max_force):
    from parity import rel_inf

    a = _run(storage, True, ib_max_force)
    a2 = _run(storage, True, ib_max_force)
    b = _run(storage, False, ib_max_force)
    active = int(a["counters"][5])
    print(f"{storage}: IB active cells {active}, clipped {int(a['counters'][6])}, total {a['ib_total']}")
    assert active > 0
    if ib_max_force is not None:
        assert int(a["counters"][6]) > 0
    tiny = 1e-12 if storage == "fp64" else 1e-6
    for key in ("post", "rho", "u", "lbm_force", "ibf", "cforce", "phi", "x", "v", "F", "pforce", "ib_total"):
        floor = rel_inf(a2[key], a[key])
        err = rel_inf(b[key], a[key])
        print(f"  {key}: two-pass vs fused {err:.3e}, fused vs fused {floor:.3e}")
        assert err <= max(4.0 * floor, tiny), (key, err, floor)
    if np.array_equal(a["counters"], a2["counters"]):
        assert np.array_equal(a["counters"], b["counters"])


def test_fused_assembly_i16_measuring_pass_counts_nothing():
    from parity import rel_inf

    ref = _run("fp32", True)
    a = _run("i16", True)
    b = _run("i16", False)
    print(f"IB active cells: fp32 {int(ref['counters'][5])}, i16 fused {int(a['counters'][5])}, "
          f"two-pass {int(b['counters'][5])}")
    assert int(a["counters"][5]) == int(ref["counters"][5]) == int(b["counters"][5]) > 0
    for key in ("rho", "u", "lbm_force", "ibf", "x", "v", "ib_total"):
        err = rel_inf(b[key], a[key])
        print(f"  {key}: two-pass vs fused {err:.3e}")
        assert err <= 1e-3, (key, err)
