"""The validation harness on the HIP path: the reference's ten validation
cases (fsi/validation_cases.py:584-596) all pass, and the six reference
datasets (fsi/reference_cases.py) compare within their tolerances against the
reference's committed data."""

from pathlib import Path

import pytest

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("gpu_available")]

CASES = ["lbm_periodic_mass_conservation", "lbm_force_response", "mpm_zero_gravity_com_stability",
         "mpm_gravity_response", "coupled_enabled_vs_disabled_drift", "coupling_force_balance",
         "coupling_force_limit", "coupling_boundary_support", "immersed_boundary_drag", "contact_diagnostics"]


def test_validation_suite_passes():
    from fsi.validation_cases import run_validation_suite

    reports = run_validation_suite()
    assert [r.case_name for r in reports] == CASES
    failed = [(r.case_name, m.name, m.value) for r in reports for m in r.metrics if not m.passed]
    assert failed == []


def test_reference_suite_matches_committed_datasets(tmp_path):
    from fsi.reference import load_reference_dataset
    from fsi.reference_cases import (default_reference_data_dir, generate_reference_datasets,
                                     run_reference_validation_suite)

    assert len(list(default_reference_data_dir().glob("*.json"))) == 6
    reports = run_reference_validation_suite()
    assert len(reports) == 6
    failed = [(r.case_name, c.metric_name, c.value, c.reference) for r in reports for c in r.comparisons
              if not c.passed]
    assert failed == []
    # regenerated datasets have the committed names and tolerances
    paths = generate_reference_datasets(tmp_path)
    assert sorted(p.name for p in paths) == sorted(p.name for p in default_reference_data_dir().glob("*.json"))
    for p in paths:
        new, old = load_reference_dataset(p), load_reference_dataset(default_reference_data_dir() / p.name)
        assert new.tolerances == old.tolerances and set(new.metrics) == set(old.metrics)


def test_snapshots_feed_the_postprocess_readers(tmp_path):
    # tests/test_immersed_boundary_contact.py:259-301 of the reference
    import numpy as np

    import fsi
    from fsi.config import BoundaryConfig, CouplingConfig, LBMConfig, MPMConfig, OutputConfig, SimulationConfig

    faces = {f: BoundaryConfig("periodic") for f in ("x_left", "x_right", "y_left", "y_right", "z_left", "z_right")}
    cfg = SimulationConfig(
        num_steps=1, lbm_dt=1.0e-3,
        lbm=LBMConfig(nx=8, ny=8, nz=8, viscosity=0.1, initial_velocity=(1.0e-3, 0.0, 0.0), **faces),
        mpm=MPMConfig(nx=8, ny=8, nz=8, max_particles=128, dt=1.0e-3, density=1.0, youngs_modulus=100.0,
                      poisson_ratio=0.25, gravity=(0.0, 0.0, 0.0), boundary_width=1),
        coupling=CouplingConfig(gamma=0.0, mpm_substeps_per_lbm_step=1, immersed_boundary_enabled=True,
                                immersed_boundary_drag=0.1, immersed_boundary_fraction_threshold=0.01,
                                contact_enabled=True, contact_velocity_damping=0.1,
                                contact_fraction_threshold=0.01),
        output=OutputConfig(output_dir=tmp_path, output_interval=1, output_format="npz"))
    solid = np.zeros((8, 8, 8), np.int32)
    solid[3:6, 3:6, 3:6] = 1
    sim = fsi.FSISimulation(cfg)
    sim.initialize_lbm(solid_np=solid)
    sim.initialize_mpm_from_numpy(np.array([[4.0, 4.0, 4.0]], np.float32), np.array([[0.1, 0.0, 0.0]], np.float32),
                                  particle_mass=1.0, particle_volume=4.0)
    sim.step()
    data = np.load(sim.write_snapshot()[0])
    ts = fsi.extract_snapshot_timeseries(tmp_path)
    assert data["immersed_boundary_force"].shape == (8, 8, 8, 3)
    assert data["particle_contact_mask"].shape == (1,)
    assert int(data["contact_candidate_count"]) >= 1
    for key in ("ib_force_norm", "ib_total_force_norm", "contact_candidate_count", "contact_damped_particle_count"):
        assert key in ts, key
    assert ts["step"].tolist() == [1]
