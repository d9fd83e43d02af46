"""Asynchronous snapshots (SimulationOutputWriter.write_snapshot_async) on the
GPU: the files equal the synchronous writer's for the same state, although
the simulation steps on while the copies and the file writes are in flight;
the particle arrays come back in input order after the particle sort."""

import sys
import types

import numpy as np
import pytest

from parity import coupled_scene

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("gpu_available")]


def _sim():
    import fsi

    cfg, solid, pos, vel = coupled_scene(storage="fp32", ib=True, contact=True)
    sim = fsi.FSISimulation(cfg)
    sim.initialize_lbm(solid)
    sim.initialize_mpm_from_numpy(pos, vel, particle_volume=0.125)
    for _ in range(3):
        sim.step()
    assert sim.mpm._permuted  # the coupled step sorts the particles
    return sim


@pytest.fixture
def evtk(monkeypatch):
    calls = []
    hl = types.ModuleType("pyevtk.hl")
    hl.imageToVTK = lambda path, origin, spacing, cellData: (calls.append((path, cellData)), path + ".vti")[1]
    hl.pointsToVTK = lambda path, x, y, z, data: (calls.append((path, (x, y, z, data))), path + ".vtu")[1]
    pkg = types.ModuleType("pyevtk")
    pkg.hl = hl
    monkeypatch.setitem(sys.modules, "pyevtk", pkg)
    monkeypatch.setitem(sys.modules, "pyevtk.hl", hl)
    return calls


def test_async_npz_snapshot_equals_the_synchronous_one(tmp_path):
    from fsi.config import OutputConfig
    from fsi.output import SimulationOutputWriter

    sim = _sim()
    sync = SimulationOutputWriter(OutputConfig(output_dir=tmp_path / "sync"))
    asyn = SimulationOutputWriter(OutputConfig(output_dir=tmp_path / "async"))
    (p_sync,) = sync.write_snapshot(sim)
    x_then = sim.mpm.positions_numpy()
    fut = asyn.write_snapshot_async(sim)
    for _ in range(2):  # the state moves on while the snapshot is in flight
        sim.step()
    assert np.abs(sim.mpm.positions_numpy() - x_then).max() > 0
    (p_async,) = fut.result(timeout=120)
    assert p_async.name == p_sync.name == "snapshot_000003.npz"
    with np.load(p_sync) as a, np.load(p_async) as b:
        assert set(a.files) == set(b.files) and len(a.files) == 29
        for k in a.files:
            assert a[k].dtype == b[k].dtype and a[k].shape == b[k].shape, k
            assert np.array_equal(a[k], b[k]), k
        assert np.array_equal(b["mpm_positions"], x_then)


def test_async_vtk_snapshot_equals_the_synchronous_one(tmp_path, evtk):
    from fsi.config import OutputConfig
    from fsi.output import SimulationOutputWriter

    sim = _sim()
    cfg = OutputConfig(output_dir=tmp_path, output_format="vtk")
    SimulationOutputWriter(cfg).write_snapshot(sim)
    sync_calls = list(evtk)
    evtk.clear()
    fut = SimulationOutputWriter(cfg).write_snapshot_async(sim)
    sim.step()
    paths = fut.result(timeout=120)
    assert [p.name for p in paths] == ["fluid_000003.vti", "particles_000003.vtu"]
    (cells_s, pts_s), (cells_a, pts_a) = [c[1] for c in sync_calls], [c[1] for c in evtk]
    assert set(cells_s) == set(cells_a)
    for k in cells_s:
        a, b = cells_s[k], cells_a[k]
        for u, v in zip(a if isinstance(a, tuple) else (a,), b if isinstance(b, tuple) else (b,)):
            assert np.array_equal(u, v), k
    for u, v in zip(pts_s[:3], pts_a[:3]):
        assert np.array_equal(u, v)
    assert set(pts_s[3]) == set(pts_a[3]) == {"velocity", "particle_force", "contact_mask"}
