"""Snapshot writer schema on the CPU (fsi/output.py of the reference).

A duck-typed simulation with the solvers' accessors (fixed arrays, no GPU)
feeds the writer; a stub ``pyevtk.hl`` module records the imageToVTK /
pointsToVTK calls (pyevtk is not installed here). Checked: the NPZ keys,
dtypes and shapes; the VTK file names, array names, spacing and origin; and
the three write_* flags, against the reference's fsi/output.py:39-196.
"""

from __future__ import annotations

import sys
import types
from types import SimpleNamespace

import numpy as np
import pytest

from fsi.config import OutputConfig
from fsi.output import SimulationOutputWriter

N = (5, 6, 7)
P = 4


def _fake_sim(dx=0.5):
    rng = np.random.default_rng(0)
    f32 = np.float32
    lbm = SimpleNamespace(
        nx=N[0], ny=N[1], nz=N[2],
        density_numpy=lambda: np.full(N, 1.0, f32),
        velocity_numpy=lambda: rng.standard_normal((*N, 3)).astype(f32),
        force_numpy=lambda: np.zeros((*N, 3), f32),
        solid_numpy=lambda: np.zeros(N, np.int32),
    )
    mpm = SimpleNamespace(
        dx=dx,
        particle_count=lambda: P,
        positions_numpy=lambda: rng.uniform(1, 4, (P, 3)).astype(f32),
        velocities_numpy=lambda: np.zeros((P, 3), f32),
        particle_forces_numpy=lambda: np.ones((P, 3), f32),
        deformation_gradients_numpy=lambda: np.broadcast_to(np.eye(3, dtype=f32), (P, 3, 3)).copy(),
        active_numpy=lambda: np.ones(P, np.int32),
    )
    diag = {"ib_total_force_x": 1.0, "ib_total_force_y": 2.0, "ib_total_force_z": 3.0,
            "ib_active_cell_count": 5, "ib_clipped_cell_count": 1, "contact_candidate_count": 2,
            "contact_damped_particle_count": 1, "unsupported_particle_count": 0,
            "partial_support_particle_count": 3, "clipped_particle_count": 0}
    coupler = SimpleNamespace(
        coupling_diagnostics=lambda: dict(diag),
        coupling_force_numpy=lambda: np.zeros((*N, 3), f32),
        solid_volume_fraction_numpy=lambda: np.zeros(N, f32),
        immersed_boundary_force_numpy=lambda: np.zeros((*N, 3), f32),
        total_particle_coupling_force=lambda: np.zeros(3, f32),
        total_fluid_coupling_force=lambda: np.zeros(3, f32),
        particle_valid_weights_numpy=lambda: np.ones(P, f32),
        particle_coupling_mask_numpy=lambda: np.ones(P, np.int32),
        particle_contact_mask_numpy=lambda: np.zeros(P, np.int32),
    )
    return SimpleNamespace(step_index=12, time=0.75, lbm=lbm, mpm=mpm, coupler=coupler)


@pytest.fixture
def evtk(monkeypatch):
    calls = []

    def image_to_vtk(path, origin=(0.0, 0.0, 0.0), spacing=(1.0, 1.0, 1.0), cellData=None, pointData=None):
        calls.append(("image", path, origin, spacing, cellData))
        return path + ".vti"

    def points_to_vtk(path, x, y, z, data=None):
        calls.append(("points", path, (x, y, z), data))
        return path + ".vtu"

    hl = types.ModuleType("pyevtk.hl")
    hl.imageToVTK, hl.pointsToVTK = image_to_vtk, points_to_vtk
    pkg = types.ModuleType("pyevtk")
    pkg.hl = hl
    monkeypatch.setitem(sys.modules, "pyevtk", pkg)
    monkeypatch.setitem(sys.modules, "pyevtk.hl", hl)
    return calls


NPZ_ALWAYS = {"step": np.int64, "time": np.float64, "lbm_shape": np.int32, "mpm_particle_count": np.int32}
NPZ_LBM = {"lbm_density", "lbm_velocity", "lbm_force", "lbm_solid"}
NPZ_MPM = {"mpm_positions", "mpm_velocities", "mpm_particle_forces", "mpm_deformation_gradients", "mpm_active"}
NPZ_COUPLING = {"coupling_force", "solid_volume_fraction", "immersed_boundary_force",
                "total_particle_coupling_force", "total_fluid_coupling_force", "coupling_particle_valid_weight",
                "coupling_particle_mask", "particle_contact_mask", "ib_total_force", "ib_active_cell_count",
                "ib_clipped_cell_count", "contact_candidate_count", "contact_damped_particle_count",
                "coupling_unsupported_particle_count", "coupling_partial_support_particle_count",
                "coupling_clipped_particle_count"}


@pytest.mark.parametrize("flags", [(True, True, True), (True, False, False), (False, True, False),
                                   (False, False, True)])
def test_npz_snapshot_keys_follow_the_flags(tmp_path, flags):
    lbm, mpm, cpl = flags
    cfg = OutputConfig(output_dir=tmp_path, output_format="npz", write_lbm_fields=lbm,
                       write_mpm_particles=mpm, write_coupling_fields=cpl)
    (path,) = SimulationOutputWriter(cfg).write_snapshot(_fake_sim())
    assert path.name == "snapshot_000012.npz"
    with np.load(path) as z:
        keys = set(z.files)
        expect = set(NPZ_ALWAYS) | (NPZ_LBM if lbm else set()) | (NPZ_MPM if mpm else set()) | \
            (NPZ_COUPLING if cpl else set())
        assert keys == expect
        for k, dt in NPZ_ALWAYS.items():
            assert z[k].dtype == dt
        assert z["step"] == 12 and z["time"] == 0.75
        assert list(z["lbm_shape"]) == list(N)
        if cpl:
            assert z["ib_total_force"].dtype == np.float32 and list(z["ib_total_force"]) == [1.0, 2.0, 3.0]
            assert z["coupling_partial_support_particle_count"] == 3
            assert z["ib_active_cell_count"].dtype == np.int32
        if mpm:
            assert z["mpm_deformation_gradients"].shape == (P, 3, 3)


def test_vtk_snapshot_names_fields_and_spacing(tmp_path, evtk):
    cfg = OutputConfig(output_dir=tmp_path, output_format="vtk")
    paths = SimulationOutputWriter(cfg).write_snapshot(_fake_sim(dx=0.5))
    assert [p.name for p in paths] == ["fluid_000012.vti", "particles_000012.vtu"]
    (kind, path, origin, spacing, cells), (kind2, path2, xyz, points) = evtk
    assert kind == "image" and path.endswith("fluid_000012") and origin == (0.0, 0.0, 0.0)
    assert spacing == (0.5, 0.5, 0.5)
    assert set(cells) == {"rho", "velocity", "force", "solid", "coupling_force", "solid_volume_fraction",
                          "immersed_boundary_force"}
    for vec in ("velocity", "force", "coupling_force", "immersed_boundary_force"):
        assert isinstance(cells[vec], tuple) and len(cells[vec]) == 3
        assert all(c.shape == N and c.flags["C_CONTIGUOUS"] for c in cells[vec])
    assert cells["rho"].shape == N
    assert kind2 == "points" and path2.endswith("particles_000012")
    assert all(c.shape == (P,) for c in xyz)
    assert set(points) == {"velocity", "particle_force", "contact_mask"}
    assert len(points["particle_force"]) == 3


def test_vtk_snapshot_flags(tmp_path, evtk):
    cfg = OutputConfig(output_dir=tmp_path, output_format="vtk", write_lbm_fields=False,
                       write_coupling_fields=True, write_mpm_particles=False)
    paths = SimulationOutputWriter(cfg).write_snapshot(_fake_sim())
    assert [p.name for p in paths] == ["fluid_000012.vti"]
    assert set(evtk[0][4]) == {"coupling_force", "solid_volume_fraction", "immersed_boundary_force"}
    evtk.clear()
    cfg = OutputConfig(output_dir=tmp_path, output_format="vtk", write_lbm_fields=False,
                       write_coupling_fields=False, write_mpm_particles=True)
    paths = SimulationOutputWriter(cfg).write_snapshot(_fake_sim())
    assert [p.name for p in paths] == ["particles_000012.vtu"]
    assert set(evtk[0][3]) == {"velocity", "particle_force"}


def test_both_formats_and_bad_format(tmp_path, evtk):
    cfg = OutputConfig(output_dir=tmp_path, output_format="both")
    paths = SimulationOutputWriter(cfg).write_snapshot(_fake_sim())
    assert [p.name for p in paths] == ["snapshot_000012.npz", "fluid_000012.vti", "particles_000012.vtu"]
    writer = SimulationOutputWriter(OutputConfig(output_dir=tmp_path))
    writer.config = SimpleNamespace(output_format="hdf5")
    with pytest.raises(ValueError, match="Unsupported output format: hdf5."):
        writer.write_snapshot(_fake_sim())
