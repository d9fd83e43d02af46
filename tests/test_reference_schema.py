"""Validation metrics and the reference dataset schema (fsi.validation,
fsi.reference) on the CPU, against the reference's committed datasets."""

import math
from pathlib import Path

import pytest

from fsi.reference import (
    ReferenceDataset,
    ReferenceMetricTolerance,
    compare_metrics,
    load_reference_dataset,
    reference_report_to_validation_report,
    save_reference_dataset,
)
from fsi.validation import bounded_metric, finite_metric, relative_error

GOLDEN = Path(__file__).resolve().parent / "golden" / "reference"


def test_bounded_metric_rules():
    assert bounded_metric("a", 0.5, lower=0.0, upper=1.0).passed
    assert not bounded_metric("a", 1.5, upper=1.0).passed
    assert not bounded_metric("a", -0.1, lower=0.0).passed
    assert not bounded_metric("a", math.nan).passed
    assert not finite_metric("a", math.inf).passed
    assert finite_metric("a", 3.0).passed
    assert relative_error(2.0, 0.0) == 2.0 and relative_error(1.1, 1.0) == pytest.approx(0.1)


@pytest.mark.parametrize("path", sorted(GOLDEN.glob("*.json")), ids=lambda p: p.stem)
def test_committed_datasets_round_trip(path, tmp_path):
    ds = load_reference_dataset(path)
    assert ds.schema_version == 1 and ds.case_name == path.stem
    assert set(ds.metrics) == set(ds.tolerances)
    again = load_reference_dataset(save_reference_dataset(ds, tmp_path / path.name))
    assert again == ds
    # the committed values compare equal to themselves
    assert compare_metrics(ds.case_name, ds.metrics, ds).passed


def test_compare_metrics_abs_or_rel():
    ds = ReferenceDataset(1, "c", "d", {"x": 2.0, "z": 0.0},
                          {"x": ReferenceMetricTolerance(abs=1e-3, rel=0.0),
                           "z": ReferenceMetricTolerance(abs=0.0, rel=0.5)})
    assert compare_metrics("c", {"x": 2.0005, "z": 0.4}, ds).passed   # z: rel error = abs error 0.4
    r = compare_metrics("c", {"x": 2.01, "z": 0.0}, ds)
    assert not r.passed and [c.passed for c in r.comparisons] == [False, True]
    v = reference_report_to_validation_report(r)
    assert v.case_name == "reference_c" and [m.name for m in v.metrics] == ["x_abs_error", "z_abs_error"]
    assert v.metrics[0].upper == pytest.approx(1e-3) and v.metrics[1].upper == pytest.approx(0.5)
    with pytest.raises(KeyError):
        compare_metrics("c", {"x": 2.0}, ds)


def test_schema_errors(tmp_path):
    bad = tmp_path / "bad.json"
    bad.write_text('{"schema_version": 2, "case_name": "c", "description": "", "metrics": {"a": 1}, "tolerances": {"a": {}}}')
    with pytest.raises(ValueError):
        load_reference_dataset(bad)
    bad.write_text('{"schema_version": 1, "case_name": "c", "description": "", "metrics": {"a": 1}, "tolerances": {}}')
    with pytest.raises(ValueError):
        load_reference_dataset(bad)
    bad.write_text('{"schema_version": 1, "case_name": "c", "description": "", "metrics": {"a": "x"}, "tolerances": {"a": {}}}')
    with pytest.raises(ValueError):
        load_reference_dataset(bad)


def test_postprocess_reads_snapshots(tmp_path):
    import numpy as np

    from fsi.postprocess import (extract_snapshot_timeseries, list_npz_snapshots, snapshot_info,
                                 summarize_coupling_diagnostics)

    rng = np.random.default_rng(0)
    for step in (10, 2, 0):
        np.savez(tmp_path / f"snapshot_{step:06d}.npz", step=np.int64(step), time=np.float64(0.1 * step),
                 mpm_particle_count=np.int64(3), lbm_shape=np.array([2, 2, 2]),
                 lbm_density=np.full((2, 2, 2), 1.0 + step, np.float32),
                 lbm_velocity=rng.normal(size=(2, 2, 2, 3)).astype(np.float32),
                 mpm_positions=np.arange(9, dtype=np.float32).reshape(3, 3),
                 ib_total_force=np.array([3.0, 4.0, 0.0]), contact_candidate_count=np.int64(step))
    assert [p.name for p in list_npz_snapshots(tmp_path)] == [
        "snapshot_000000.npz", "snapshot_000002.npz", "snapshot_000010.npz"]
    assert snapshot_info(list_npz_snapshots(tmp_path)[1]).step == 2
    ts = extract_snapshot_timeseries(tmp_path)
    assert list(ts)[:3] == ["step", "time", "mpm_particle_count"]
    assert ts["step"].tolist() == [0, 2, 10]
    assert ts["lbm_total_mass_estimate"].tolist() == [8.0, 24.0, 88.0]
    assert ts["mpm_center_of_mass_y"].tolist() == [4.0, 4.0, 4.0]
    assert ts["ib_total_force_norm"].tolist() == [5.0, 5.0, 5.0]
    assert "contact_candidate_count" in summarize_coupling_diagnostics(tmp_path)


def test_package_exports_resolve():
    import fsi

    for name in fsi.__all__:
        if name in ("LBMSolver3D", "MPMSolver3D", "LBMMpmCoupler", "FSISimulation", "SimulationOutputWriter",
                    "run_validation_suite", "run_reference_validation_suite"):
            continue  # solver modules: resolved by the GPU tests
        assert getattr(fsi, name) is not None, name


def test_packaged_datasets_are_the_test_fixtures():
    """The validation harness's datasets (package data) are byte for byte the
    reference's committed files kept as test fixtures."""
    from pathlib import Path

    from fsi.reference_cases import default_reference_data_dir

    pkg = default_reference_data_dir()
    golden = Path(__file__).resolve().parent / "golden" / "reference"
    names = sorted(p.name for p in golden.glob("*.json"))
    assert len(names) == 6 and sorted(p.name for p in pkg.glob("*.json")) == names
    for n in names:
        assert (pkg / n).read_bytes() == (golden / n).read_bytes(), n
