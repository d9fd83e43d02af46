"""The reference's ten small validation scenarios, written against an API namespace.

Each ``case_*`` function rebuilds one scenario of the reference's
``fsi/validation_cases.py`` (cited per function) through whatever object
factory it is given:

* ``oracle_api(precision)`` - the CPU oracle classes (test infrastructure);
* ``product_api()``         - the ``fsi`` package (HIP kernels on the GPU).

They return the scalar metrics the reference's fixtures in
``data/reference/*.json`` were generated from (``fsi/reference_cases.py:35-84``),
so the same code checks both the oracle and the product against the fixtures.
"""

from __future__ import annotations

import json
from dataclasses import dataclass
from pathlib import Path
from types import SimpleNamespace

import numpy as np

from fsi.config import BoundaryConfig, CouplingConfig, LBMConfig, MPMConfig, SimulationConfig

GOLDEN = Path(__file__).resolve().parent / "golden"


def oracle_api(precision: str = "f32") -> SimpleNamespace:
    from oracle import pyoracle as po

    return SimpleNamespace(
        name=f"oracle-{precision}",
        LBMSolver3D=lambda cfg: po.OracleLBM(cfg, precision),
        MPMSolver3D=lambda cfg: po.OracleMPM(cfg, precision),
        LBMMpmCoupler=po.OracleCoupler,
        FSISimulation=lambda cfg: po.OracleSimulation(cfg, precision),
    )


def product_api(storage: str = "fp32") -> SimpleNamespace:
    import dataclasses

    import fsi

    def lbm_cfg(cfg):
        return dataclasses.replace(cfg, storage=storage)

    def sim_cfg(cfg):
        return dataclasses.replace(cfg, lbm=lbm_cfg(cfg.lbm))

    return SimpleNamespace(
        name=f"hip-{storage}",
        LBMSolver3D=lambda cfg: fsi.LBMSolver3D(lbm_cfg(cfg)),
        MPMSolver3D=fsi.MPMSolver3D,
        LBMMpmCoupler=fsi.LBMMpmCoupler,
        FSISimulation=lambda cfg: fsi.FSISimulation(sim_cfg(cfg)),
    )


def periodic_lbm_config(**overrides) -> LBMConfig:
    # fsi/validation_cases.py:12-29
    params = dict(nx=8, ny=8, nz=8, viscosity=0.1, rho0=1.0, force=(0.0, 0.0, 0.0),
                  initial_velocity=(0.0, 0.0, 0.0))
    for face in ("x_left", "x_right", "y_left", "y_right", "z_left", "z_right"):
        params[face] = BoundaryConfig("periodic")
    params.update(overrides)
    return LBMConfig(**params)


def small_mpm_config(**overrides) -> MPMConfig:
    # fsi/validation_cases.py:32-47
    params = dict(nx=16, ny=16, nz=16, max_particles=10_000, dx=1.0, dt=1.0e-3, density=1.0,
                  youngs_modulus=1_000.0, poisson_ratio=0.25, gravity=(0.0, 0.0, 0.0),
                  boundary_width=2)
    params.update(overrides)
    return MPMConfig(**params)


def coupled_mpm_config(**overrides) -> MPMConfig:
    # the 8^3 MPM block used by every coupled scenario (validation_cases.py:54-66)
    params = dict(nx=8, ny=8, nz=8, max_particles=128, dx=1.0, dt=1.0e-3, density=1.0,
                  youngs_modulus=100.0, poisson_ratio=0.25, gravity=(0.0, 0.0, 0.0),
                  boundary_width=1)
    params.update(overrides)
    return MPMConfig(**params)


def one_particle_sim(api, config, position=(4.0, 4.0, 4.0), velocity=(0.0, 0.0, 0.0),
                     volume=1.0, solid=None, lbm_first=False):
    # validation_cases.py:71-79 (MPM first) and the IB/contact variants (LBM first)
    sim = api.FSISimulation(config)
    if lbm_first:
        sim.initialize_lbm(solid_np=solid)
    sim.initialize_mpm_from_numpy(
        positions=np.array([position], dtype=np.float32),
        velocities=np.array([velocity], dtype=np.float32),
        particle_mass=1.0,
        particle_volume=volume,
    )
    if not lbm_first:
        sim.initialize_lbm(solid_np=solid)
    return sim


def case_lbm_mass(api):
    # validation_cases.py:82-114
    solver = api.LBMSolver3D(periodic_lbm_config())
    solver.initialize()
    m0 = solver.total_mass()
    for _ in range(20):
        solver.step()
    m1 = solver.total_mass()
    rel = abs((m1 - m0) / m0) if m0 != 0 else abs(m1 - m0)
    return {"relative_mass_error": float(rel), "max_velocity_norm": solver.max_velocity_norm()}


def case_lbm_force_response(api):
    # validation_cases.py:117-146
    solver = api.LBMSolver3D(periodic_lbm_config(force=(1.0e-6, 0.0, 0.0)))
    solver.initialize()
    ux0 = float(solver.velocity_numpy()[..., 0].mean())
    for _ in range(20):
        solver.step()
    ux1 = float(solver.velocity_numpy()[..., 0].mean())
    return {"mean_ux_growth": ux1 - ux0, "max_velocity_norm": solver.max_velocity_norm()}


def case_mpm_zero_gravity(api):
    # validation_cases.py:148-176
    solver = api.MPMSolver3D(small_mpm_config(gravity=(0.0, 0.0, 0.0), dt=1.0e-3))
    solver.initialize_particles_box((6.0, 6.0, 6.0), (10.0, 10.0, 10.0), spacing=1.0)
    c0 = solver.center_of_mass()
    for _ in range(10):
        solver.substep()
    c1 = solver.center_of_mass()
    return {"center_of_mass_drift_norm": float(np.linalg.norm(c1 - c0)),
            "max_velocity_norm": solver.max_velocity_norm()}


def case_mpm_gravity(api):
    # validation_cases.py:179-218
    solver = api.MPMSolver3D(small_mpm_config(gravity=(0.0, -1.0e-3, 0.0), dt=1.0e-2))
    solver.initialize_particles_box((6.0, 8.0, 6.0), (10.0, 12.0, 10.0), spacing=1.0)
    c0 = solver.center_of_mass()
    for _ in range(10):
        solver.substep()
    c1 = solver.center_of_mass()
    return {
        "center_of_mass_y_delta": float(c1[1] - c0[1]),
        "positions_finite": 1.0 if np.isfinite(solver.positions_numpy()).all() else 0.0,
        "velocities_finite": 1.0 if np.isfinite(solver.velocities_numpy()).all() else 0.0,
    }


def _coupled_drift_config(enabled):
    # validation_cases.py:50-68
    return SimulationConfig(
        num_steps=2, lbm_dt=0.1,
        lbm=periodic_lbm_config(initial_velocity=(1.0e-3, 0.0, 0.0)),
        mpm=coupled_mpm_config(),
        coupling=CouplingConfig(enabled=enabled, gamma=5.0, mpm_substeps_per_lbm_step=1),
    )


def case_coupled_drift(api):
    # validation_cases.py:220-260
    on = one_particle_sim(api, _coupled_drift_config(True))
    off = one_particle_sim(api, _coupled_drift_config(False))
    x_on = float(on.mpm.center_of_mass()[0])
    x_off = float(off.mpm.center_of_mass()[0])
    on.run(steps=2)
    off.run(steps=2)
    dx_on = float(on.mpm.center_of_mass()[0] - x_on)
    dx_off = float(off.mpm.center_of_mass()[0] - x_off)
    return {"enabled_dx": dx_on, "disabled_dx_abs": abs(dx_off),
            "enabled_minus_disabled_dx": dx_on - dx_off}


def _balance(sim):
    pf = sim.coupler.total_particle_coupling_force()
    ff = sim.coupler.total_fluid_coupling_force() * sim.coupler.cell_volume
    return pf, ff, float(np.linalg.norm(pf + ff))


def case_force_balance(api):
    # validation_cases.py:262-297
    sim = one_particle_sim(api, _coupled_drift_config(True))
    sim.coupler.compute_coupling_forces(dt_ratio=1.0)
    pf, ff, bal = _balance(sim)
    return {"particle_force_x": float(pf[0]), "fluid_force_x": float(ff[0]),
            "force_balance_norm": bal}


def case_force_limit(api):
    # validation_cases.py:300-352
    cfg = SimulationConfig(
        num_steps=1, lbm_dt=0.1, lbm=periodic_lbm_config(initial_velocity=(1.0, 0.0, 0.0)),
        mpm=coupled_mpm_config(),
        coupling=CouplingConfig(gamma=100.0, force_limit=0.05, mpm_substeps_per_lbm_step=1),
    )
    sim = one_particle_sim(api, cfg)
    sim.coupler.compute_coupling_forces(dt_ratio=1.0)
    pf, ff, bal = _balance(sim)
    d = sim.coupler.coupling_diagnostics()
    return {"particle_force_norm": float(np.linalg.norm(pf)),
            "clipped_particle_count": float(d["clipped_particle_count"]),
            "force_balance_norm": bal}


def case_boundary_support(api):
    # validation_cases.py:355-426
    cfg = SimulationConfig(
        num_steps=1, lbm_dt=0.1, lbm=periodic_lbm_config(initial_velocity=(1.0, 0.0, 0.0)),
        mpm=coupled_mpm_config(),
        coupling=CouplingConfig(gamma=1.0, min_valid_weight=1.0e-6, mpm_substeps_per_lbm_step=1),
    )
    sim = one_particle_sim(api, cfg, position=(-0.75, 4.0, 4.0))
    sim.coupler.compute_coupling_forces(dt_ratio=1.0)
    pf, ff, bal = _balance(sim)
    d = sim.coupler.coupling_diagnostics()
    return {"partial_support_particle_count": float(d["partial_support_particle_count"]),
            "min_particle_valid_weight": float(d["min_particle_valid_weight"]),
            "force_balance_norm": bal,
            "particle_force_finite": 1.0 if np.isfinite(pf).all() else 0.0}


def case_ib_drag(api):
    # validation_cases.py:428-506
    cfg = SimulationConfig(
        num_steps=1, lbm_dt=1.0e-3, lbm=periodic_lbm_config(initial_velocity=(1.0e-3, 0.0, 0.0)),
        mpm=coupled_mpm_config(),
        coupling=CouplingConfig(gamma=0.0, mpm_substeps_per_lbm_step=1,
                                immersed_boundary_enabled=True, immersed_boundary_drag=0.1,
                                immersed_boundary_fraction_threshold=0.01,
                                immersed_boundary_max_force=1.0e-3),
    )
    sim = one_particle_sim(api, cfg, volume=4.0, lbm_first=True)
    s0 = sim.lbm.max_velocity_norm()
    sim.step()
    d = sim.coupler.coupling_diagnostics()
    return {"ib_active_cell_count": float(d["ib_active_cell_count"]),
            "ib_total_force_x": float(d["ib_total_force_x"]),
            "ib_total_force_norm": float(d["ib_total_force_norm"]),
            "max_velocity_not_increased": float(sim.lbm.max_velocity_norm() - s0),
            "velocity_finite": 1.0 if np.isfinite(sim.lbm.velocity_numpy()).all() else 0.0}


def case_contact(api):
    # validation_cases.py:509-581
    cfg = SimulationConfig(
        num_steps=1, lbm_dt=1.0e-3, lbm=periodic_lbm_config(),
        mpm=coupled_mpm_config(),
        coupling=CouplingConfig(gamma=0.0, mpm_substeps_per_lbm_step=1, contact_enabled=True,
                                contact_velocity_damping=0.25, contact_fraction_threshold=0.01),
    )
    solid = np.zeros((8, 8, 8), np.int32)
    solid[3:6, 3:6, 3:6] = 1
    sim = one_particle_sim(api, cfg, velocity=(0.1, 0.0, 0.0), solid=solid, lbm_first=True)
    s0 = sim.mpm.max_velocity_norm()
    sim.step()
    d = sim.coupler.coupling_diagnostics()
    return {"contact_candidate_count": float(d["contact_candidate_count"]),
            "contact_damped_particle_count": float(d["contact_damped_particle_count"]),
            "particle_contact_mask_sum": float(sim.coupler.particle_contact_mask_numpy().sum()),
            "particle_speed_delta": float(sim.mpm.max_velocity_norm() - s0)}


# fixture file -> (cases, metric names) as fsi/reference_cases.py:35-84 selects them
FIXTURE_CASES = {
    "lbm_periodic_mass_reference": [case_lbm_mass],
    "lbm_force_response_reference": [case_lbm_force_response],
    "mpm_gravity_reference": [case_mpm_gravity],
    "coupled_drift_reference": [case_coupled_drift],
    "coupling_stability_reference": [case_force_limit, case_boundary_support],
    "immersed_boundary_contact_reference": [case_ib_drag, case_contact],
}


def fixture_metrics(api, name: str) -> dict:
    out = {}
    for fn in FIXTURE_CASES[name]:
        out.update(fn(api))
    return out


def load_fixture(name: str) -> dict:
    return json.loads((GOLDEN / "reference" / f"{name}.json").read_text())


def compare_fixture(name: str, metrics: dict) -> list[str]:
    """fsi/reference.py:158-199: a metric passes when abs_err <= abs_tol or rel_err <= rel_tol."""
    data = load_fixture(name)
    failures = []
    for key, ref in data["metrics"].items():
        tol = data["tolerances"][key]
        value = float(metrics[key])
        ref = float(ref)
        abs_err = abs(value - ref)
        rel_err = abs_err if ref == 0.0 else abs_err / abs(ref)
        ok = abs_err <= float(tol["abs"]) or rel_err <= float(tol["rel"])
        if not (np.isfinite(value) and ok):
            failures.append(f"{name}.{key}: got {value!r}, fixture {ref!r}, tol {tol}")
    return failures
