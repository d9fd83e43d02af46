"""Reference datasets built from the validation cases, and the comparison
suite (reference: fsi/reference_cases.py).

The six datasets select metrics of the validation cases with the reference's
case names, metric names and tolerances (fsi/reference_cases.py:36-180), so
`run_reference_validation_suite()` compares this package's results with the
reference's committed datasets directly.
"""

from __future__ import annotations

from pathlib import Path

from .reference import (
    ReferenceDataset,
    ReferenceMetricTolerance,
    ReferenceReport,
    compare_metrics,
    load_reference_dataset,
    save_reference_dataset,
)
from .validation import ValidationReport
from .validation_cases import (
    run_contact_diagnostics_case,
    run_coupled_drift_case,
    run_coupling_boundary_support_case,
    run_coupling_force_limit_case,
    run_immersed_boundary_drag_case,
    run_lbm_force_response_case,
    run_lbm_mass_conservation_case,
    run_mpm_gravity_response_case,
)

CREATED_BY = "fsi MI355X reference generator"


def default_reference_data_dir() -> Path:
    """The reference's committed datasets (its data/reference/*.json), shipped
    as package data of fsi."""
    return Path(__file__).resolve().parent / "data" / "reference"


def _selected(report: ValidationReport, names: tuple[str, ...]) -> dict[str, float]:
    values = {m.name: float(m.value) for m in report.metrics}
    return {n: values[n] for n in names}


def compute_lbm_periodic_mass_reference_metrics() -> dict[str, float]:
    return _selected(run_lbm_mass_conservation_case(), ("relative_mass_error", "max_velocity_norm"))


def compute_lbm_force_response_reference_metrics() -> dict[str, float]:
    return _selected(run_lbm_force_response_case(), ("mean_ux_growth", "max_velocity_norm"))


def compute_mpm_gravity_reference_metrics() -> dict[str, float]:
    return _selected(run_mpm_gravity_response_case(),
                     ("center_of_mass_y_delta", "positions_finite", "velocities_finite"))


def compute_coupled_drift_reference_metrics() -> dict[str, float]:
    return _selected(run_coupled_drift_case(), ("enabled_dx", "disabled_dx_abs", "enabled_minus_disabled_dx"))


def compute_coupling_stability_reference_metrics() -> dict[str, float]:
    return {**_selected(run_coupling_force_limit_case(), ("particle_force_norm", "force_balance_norm")),
            **_selected(run_coupling_boundary_support_case(),
                        ("partial_support_particle_count", "min_particle_valid_weight"))}


def compute_immersed_boundary_contact_reference_metrics() -> dict[str, float]:
    return {**_selected(run_immersed_boundary_drag_case(),
                        ("ib_active_cell_count", "ib_total_force_x", "ib_total_force_norm")),
            **_selected(run_contact_diagnostics_case(),
                        ("contact_candidate_count", "contact_damped_particle_count"))}


# case name -> (description, metric function, tolerances {metric: (abs, rel)}, metadata);
# the tolerances are the reference's (fsi/reference_cases.py:87-180)
_DATASETS = {
    "lbm_periodic_mass_reference": (
        "Small periodic LBM mass conservation reference.", compute_lbm_periodic_mass_reference_metrics,
        {"relative_mass_error": (1.0e-5, 0.0), "max_velocity_norm": (1.0e-4, 0.0)},
        {"source_case": "lbm_periodic_mass_conservation", "steps": 20}),
    "lbm_force_response_reference": (
        "Small forced periodic LBM velocity-growth reference.", compute_lbm_force_response_reference_metrics,
        {"mean_ux_growth": (1.0e-8, 1.0e-2), "max_velocity_norm": (1.0e-8, 1.0e-2)},
        {"source_case": "lbm_force_response", "steps": 20}),
    "mpm_gravity_reference": (
        "Weak-gravity MPM response reference.", compute_mpm_gravity_reference_metrics,
        {"center_of_mass_y_delta": (1.0e-6, 1.0e-2), "positions_finite": (0.0, 0.0),
         "velocities_finite": (0.0, 0.0)},
        {"source_case": "mpm_gravity_response", "steps": 10}),
    "coupled_drift_reference": (
        "Enabled-vs-disabled coupling drift reference.", compute_coupled_drift_reference_metrics,
        {"enabled_dx": (1.0e-6, 5.0e-2), "disabled_dx_abs": (1.0e-6, 0.0),
         "enabled_minus_disabled_dx": (1.0e-6, 5.0e-2)},
        {"source_case": "coupled_enabled_vs_disabled_drift", "steps": 2}),
    "coupling_stability_reference": (
        "Force-limit and boundary-support reference.", compute_coupling_stability_reference_metrics,
        {"particle_force_norm": (1.0e-7, 1.0e-3), "force_balance_norm": (1.0e-6, 0.0),
         "partial_support_particle_count": (0.0, 0.0), "min_particle_valid_weight": (1.0e-6, 1.0e-3)},
        {"source_cases": ["coupling_force_limit", "coupling_boundary_support"]}),
    "immersed_boundary_contact_reference": (
        "Small immersed-boundary/contact reference.", compute_immersed_boundary_contact_reference_metrics,
        {"ib_active_cell_count": (0.0, 0.0), "ib_total_force_x": (1.0e-9, 1.0e-3),
         "ib_total_force_norm": (1.0e-9, 1.0e-3), "contact_candidate_count": (0.0, 0.0),
         "contact_damped_particle_count": (0.0, 0.0)},
        {"source_cases": ["immersed_boundary_drag", "contact_diagnostics"]}),
}


def build_reference_datasets() -> list[ReferenceDataset]:
    """The six datasets computed from the current implementation."""
    return [ReferenceDataset(schema_version=1, case_name=name, description=desc, metrics=fn(),
                             tolerances={k: ReferenceMetricTolerance(abs=a, rel=r) for k, (a, r) in tol.items()},
                             metadata=meta, created_by=CREATED_BY)
            for name, (desc, fn, tol, meta) in _DATASETS.items()]


def generate_reference_datasets(output_dir: str | Path) -> list[Path]:
    """Write freshly computed datasets (`<case_name>.json`) to `output_dir`."""
    return [save_reference_dataset(d, Path(output_dir) / f"{d.case_name}.json")
            for d in build_reference_datasets()]


def run_reference_validation_suite(reference_dir: str | Path | None = None) -> list[ReferenceReport]:
    """Compare the current metrics with every dataset in `reference_dir`
    (default: the committed datasets), in file-name order."""
    directory = default_reference_data_dir() if reference_dir is None else Path(reference_dir)
    current = {name: spec[1]() for name, spec in _DATASETS.items()}
    return [compare_metrics(d.case_name, current[d.case_name], d)
            for d in (load_reference_dataset(p) for p in sorted(directory.glob("*.json")))]
