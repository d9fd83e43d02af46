"""The reference's committed fixtures (data/reference/*.json) reproduced by the HIP path.

Each fixture is regenerated with the reference's own scenario definitions
(cases.py, from fsi/validation_cases.py) through the GPU solvers and compared
with the reference's comparison rule (fsi/reference.py:158-199).
"""

import pytest

import cases

pytestmark = [pytest.mark.gpu, pytest.mark.usefixtures("gpu_available")]


@pytest.mark.parametrize("name", list(cases.FIXTURE_CASES))
def test_fixture_reproduced_on_gpu(name):
    metrics = cases.fixture_metrics(cases.product_api("fp32"), name)
    print(name, metrics)
    assert cases.compare_fixture(name, metrics) == []


@pytest.mark.parametrize(
    "case",
    [cases.case_mpm_zero_gravity, cases.case_force_balance, cases.case_coupled_drift],
)
def test_validation_bounds_on_gpu(case):
    """Pass bounds of the scenarios without fixtures (fsi/validation_cases.py)."""
    m = case(cases.product_api("fp32"))
    if case is cases.case_mpm_zero_gravity:
        assert m["center_of_mass_drift_norm"] <= 1e-3
        assert m["max_velocity_norm"] <= 1e-3
    elif case is cases.case_force_balance:
        assert m["particle_force_x"] >= 1e-12
        assert m["fluid_force_x"] <= -1e-12
        assert m["force_balance_norm"] <= 1e-6
    else:
        assert m["enabled_dx"] >= 1e-7
        assert m["disabled_dx_abs"] <= 1e-6
