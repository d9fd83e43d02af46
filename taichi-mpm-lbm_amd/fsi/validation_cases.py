"""The reference's ten small validation scenarios on the GPU solvers
(reference: fsi/validation_cases.py).

Each `run_*_case()` builds the scene of the same-named reference function
(cited per function), runs it through this package's solvers and returns a
ValidationReport with the reference's metric names and pass bounds, so the
reports and the reference datasets built from them (fsi/reference_cases.py)
read the same.
"""

from __future__ import annotations

import numpy as np

from .config import BoundaryConfig, CouplingConfig, LBMConfig, MPMConfig, SimulationConfig
from .lbm3d import LBMSolver3D
from .mpm3d import MPMSolver3D
from .simulation import FSISimulation
from .validation import ValidationReport, bounded_metric, finite_metric, relative_error

_FACES = ("x_left", "x_right", "y_left", "y_right", "z_left", "z_right")


def _periodic_lbm_config(**overrides) -> LBMConfig:
    # fsi/validation_cases.py:12-29: 8^3, nu 0.1, rho0 1, all faces periodic
    params = dict(nx=8, ny=8, nz=8, viscosity=0.1, rho0=1.0, force=(0.0, 0.0, 0.0),
                  initial_velocity=(0.0, 0.0, 0.0), **{f: BoundaryConfig("periodic") for f in _FACES})
    params.update(overrides)
    return LBMConfig(**params)


def _mpm_config(**overrides) -> MPMConfig:
    # fsi/validation_cases.py:32-47: 16^3 box
    params = dict(nx=16, ny=16, nz=16, max_particles=10_000, dx=1.0, dt=1.0e-3, density=1.0,
                  youngs_modulus=1_000.0, poisson_ratio=0.25, gravity=(0.0, 0.0, 0.0), boundary_width=2)
    params.update(overrides)
    return MPMConfig(**params)


def _small_mpm_config() -> MPMConfig:
    # the 8^3 MPM grid of the coupled cases (fsi/validation_cases.py:53-66 and the
    # configs of 296-583)
    return MPMConfig(nx=8, ny=8, nz=8, max_particles=128, dx=1.0, dt=1.0e-3, density=1.0,
                     youngs_modulus=100.0, poisson_ratio=0.25, gravity=(0.0, 0.0, 0.0), boundary_width=1)


def _coupled_config(*, enabled: bool = True) -> SimulationConfig:
    # fsi/validation_cases.py:50-73
    return SimulationConfig(num_steps=2, lbm_dt=0.1, lbm=_periodic_lbm_config(initial_velocity=(1.0e-3, 0.0, 0.0)),
                            mpm=_small_mpm_config(),
                            coupling=CouplingConfig(enabled=enabled, gamma=5.0, mpm_substeps_per_lbm_step=1))


def _one_particle(sim: FSISimulation, position=(4.0, 4.0, 4.0), velocity=(0.0, 0.0, 0.0),
                  volume: float = 1.0) -> FSISimulation:
    sim.initialize_mpm_from_numpy(positions=np.array([position], dtype=np.float32),
                                  velocities=np.array([velocity], dtype=np.float32),
                                  particle_mass=1.0, particle_volume=volume)
    return sim


def _initialized_one_particle_simulation(config: SimulationConfig) -> FSISimulation:
    # fsi/validation_cases.py:76-85 (particles first, then the lattice)
    sim = _one_particle(FSISimulation(config))
    sim.initialize_lbm()
    return sim


def _finite_flag(name: str, arr: np.ndarray, description: str):
    return bounded_metric(name, 1.0 if np.isfinite(arr).all() else 0.0, lower=1.0, upper=1.0,
                          description=description)


def _balance(sim: FSISimulation) -> tuple[np.ndarray, np.ndarray, float]:
    particle = sim.coupler.total_particle_coupling_force()
    fluid = sim.coupler.total_fluid_coupling_force() * sim.coupler.cell_volume
    return particle, fluid, float(np.linalg.norm(particle + fluid))


def run_lbm_mass_conservation_case() -> ValidationReport:
    # fsi/validation_cases.py:88-115
    steps = 20
    s = LBMSolver3D(_periodic_lbm_config(force=(0.0, 0.0, 0.0)))
    s.initialize()
    m0 = s.total_mass()
    for _ in range(steps):
        s.step()
    m1 = s.total_mass()
    return ValidationReport(
        case_name="lbm_periodic_mass_conservation",
        metrics=(
            bounded_metric("relative_mass_error", relative_error(m1, m0), upper=1.0e-5,
                           description="Fluid mass change over the run, relative to the start (all faces periodic)."),
            finite_metric("max_velocity_norm", s.max_velocity_norm(), upper=1.0e-4,
                          description="Largest |u| stays finite and close to rest."),
        ),
        metadata={"steps": steps, "grid": [s.nx, s.ny, s.nz]})


def run_lbm_force_response_case() -> ValidationReport:
    # fsi/validation_cases.py:118-145
    steps = 20
    s = LBMSolver3D(_periodic_lbm_config(force=(1.0e-6, 0.0, 0.0)))
    s.initialize()
    ux0 = float(s.velocity_numpy()[..., 0].mean())
    for _ in range(steps):
        s.step()
    ux1 = float(s.velocity_numpy()[..., 0].mean())
    return ValidationReport(
        case_name="lbm_force_response",
        metrics=(
            bounded_metric("mean_ux_growth", ux1 - ux0, lower=1.0e-12,
                           description="A +x body force raises the mean streamwise velocity."),
            finite_metric("max_velocity_norm", s.max_velocity_norm(), upper=1.0e-2,
                          description="Largest |u| of the forced lattice stays small."),
        ),
        metadata={"steps": steps, "grid": [s.nx, s.ny, s.nz]})


def run_mpm_zero_gravity_com_case() -> ValidationReport:
    # fsi/validation_cases.py:148-176
    steps = 10
    s = MPMSolver3D(_mpm_config(gravity=(0.0, 0.0, 0.0), dt=1.0e-3))
    s.initialize_particles_box((6.0, 6.0, 6.0), (10.0, 10.0, 10.0), spacing=1.0)
    c0 = s.center_of_mass()
    for _ in range(steps):
        s.substep()
    c1 = s.center_of_mass()
    return ValidationReport(
        case_name="mpm_zero_gravity_com_stability",
        metrics=(
            bounded_metric("center_of_mass_drift_norm", float(np.linalg.norm(c1 - c0)), upper=1.0e-3,
                           units="grid", description="Centre of mass stays put when nothing drives the body."),
            finite_metric("max_velocity_norm", s.max_velocity_norm(), upper=1.0e-3,
                          description="Largest particle speed stays small with gravity off."),
        ),
        metadata={"steps": steps, "particle_count": s.particle_count()})


def run_mpm_gravity_response_case() -> ValidationReport:
    # fsi/validation_cases.py:179-217
    steps = 10
    s = MPMSolver3D(_mpm_config(gravity=(0.0, -1.0e-3, 0.0), dt=1.0e-2))
    s.initialize_particles_box((6.0, 8.0, 6.0), (10.0, 12.0, 10.0), spacing=1.0)
    c0 = s.center_of_mass()
    for _ in range(steps):
        s.substep()
    c1 = s.center_of_mass()
    return ValidationReport(
        case_name="mpm_gravity_response",
        metrics=(
            bounded_metric("center_of_mass_y_delta", float(c1[1] - c0[1]), upper=-1.0e-12, units="grid",
                           description="A small -y gravity moves the centre of mass down."),
            _finite_flag("positions_finite", s.positions_numpy(), "No particle position is NaN or infinite."),
            _finite_flag("velocities_finite", s.velocities_numpy(), "No particle velocity is NaN or infinite."),
        ),
        metadata={"steps": steps, "particle_count": s.particle_count()})


def run_coupled_drift_case() -> ValidationReport:
    # fsi/validation_cases.py:220-259
    dx = {}
    for enabled in (True, False):
        sim = _initialized_one_particle_simulation(_coupled_config(enabled=enabled))
        x0 = float(sim.mpm.center_of_mass()[0])
        sim.run(steps=2)
        dx[enabled] = float(sim.mpm.center_of_mass()[0] - x0)
    return ValidationReport(
        case_name="coupled_enabled_vs_disabled_drift",
        metrics=(
            bounded_metric("enabled_dx", dx[True], lower=1.0e-7, units="grid",
                           description="With coupling on, the flow carries the particle along +x."),
            bounded_metric("disabled_dx_abs", abs(dx[False]), upper=1.0e-6, units="grid",
                           description="With coupling off, the particle barely moves."),
            bounded_metric("enabled_minus_disabled_dx", dx[True] - dx[False], lower=1.0e-7, units="grid",
                           description="The coupled particle travels farther than the uncoupled one."),
        ),
        metadata={"steps": 2, "gamma": 5.0, "initial_lbm_ux": 1.0e-3})


def run_coupling_force_balance_case() -> ValidationReport:
    # fsi/validation_cases.py:262-294
    sim = _initialized_one_particle_simulation(_coupled_config(enabled=True))
    sim.coupler.compute_coupling_forces(dt_ratio=1.0)
    particle, fluid, norm = _balance(sim)
    return ValidationReport(
        case_name="coupling_force_balance",
        metrics=(
            bounded_metric("particle_force_x", float(particle[0]), lower=1.0e-12,
                           description="A +x stream pushes the particle along +x."),
            bounded_metric("fluid_force_x", float(fluid[0]), upper=-1.0e-12,
                           description="The lattice gets the equal and opposite x force."),
            bounded_metric("force_balance_norm", norm, upper=1.0e-6,
                           description="Particle force and summed fluid reaction cancel."),
        ),
        metadata={"dt_ratio": 1.0, "cell_volume": sim.coupler.cell_volume})


def _unit_flow_config(coupling: CouplingConfig) -> SimulationConfig:
    # fsi/validation_cases.py:297-320 and 355-378: u = (1, 0, 0), one substep
    return SimulationConfig(num_steps=1, lbm_dt=0.1, lbm=_periodic_lbm_config(initial_velocity=(1.0, 0.0, 0.0)),
                            mpm=_small_mpm_config(), coupling=coupling)


def run_coupling_force_limit_case() -> ValidationReport:
    # fsi/validation_cases.py:297-352
    sim = _initialized_one_particle_simulation(
        _unit_flow_config(CouplingConfig(gamma=100.0, force_limit=0.05, mpm_substeps_per_lbm_step=1)))
    sim.coupler.compute_coupling_forces(dt_ratio=1.0)
    particle, _, norm = _balance(sim)
    d = sim.coupler.coupling_diagnostics()
    return ValidationReport(
        case_name="coupling_force_limit",
        metrics=(
            bounded_metric("particle_force_norm", float(np.linalg.norm(particle)), upper=0.05 + 1.0e-7,
                           description="A very stiff penalty rate hits the force limit."),
            bounded_metric("clipped_particle_count", float(d["clipped_particle_count"]), lower=1.0,
                           description="The clipped-particle counter is non-zero."),
            bounded_metric("force_balance_norm", norm, upper=1.0e-6,
                           description="After clipping, particle force and fluid reaction still cancel."),
        ),
        metadata={"gamma": 100.0, "force_limit": 0.05})


def run_coupling_boundary_support_case() -> ValidationReport:
    # fsi/validation_cases.py:355-425: a particle at x = -0.75
    sim = _one_particle(FSISimulation(_unit_flow_config(
        CouplingConfig(gamma=1.0, min_valid_weight=1.0e-6, mpm_substeps_per_lbm_step=1))),
        position=(-0.75, 4.0, 4.0))
    sim.initialize_lbm()
    sim.coupler.compute_coupling_forces(dt_ratio=1.0)
    particle, _, norm = _balance(sim)
    d = sim.coupler.coupling_diagnostics()
    return ValidationReport(
        case_name="coupling_boundary_support",
        metrics=(
            bounded_metric("partial_support_particle_count", float(d["partial_support_particle_count"]),
                           lower=1.0, description="A particle next to the wall counts as partially supported."),
            bounded_metric("min_particle_valid_weight", float(d["min_particle_valid_weight"]), lower=1.0e-6,
                           upper=1.0 - 1.0e-5, description="Its fluid-node weight sum is finite and short of one."),
            bounded_metric("force_balance_norm", norm, upper=1.0e-6,
                           description="With partial support, particle force and fluid reaction still cancel."),
            _finite_flag("particle_force_finite", particle, "The wall-side particle's force is not NaN or infinite."),
        ),
        metadata={"particle_position": [-0.75, 4.0, 4.0], "min_valid_weight": 1.0e-6})


def run_immersed_boundary_drag_case() -> ValidationReport:
    # fsi/validation_cases.py:428-505
    cfg = SimulationConfig(
        num_steps=1, lbm_dt=1.0e-3, lbm=_periodic_lbm_config(initial_velocity=(1.0e-3, 0.0, 0.0)),
        mpm=_small_mpm_config(),
        coupling=CouplingConfig(gamma=0.0, mpm_substeps_per_lbm_step=1, immersed_boundary_enabled=True,
                                immersed_boundary_drag=0.1, immersed_boundary_fraction_threshold=0.01,
                                immersed_boundary_max_force=1.0e-3))
    sim = FSISimulation(cfg)
    sim.initialize_lbm()
    _one_particle(sim, volume=4.0)
    speed0 = sim.lbm.max_velocity_norm()
    sim.step()
    d = sim.coupler.coupling_diagnostics()
    return ValidationReport(
        case_name="immersed_boundary_drag",
        metrics=(
            bounded_metric("ib_active_cell_count", float(d["ib_active_cell_count"]), lower=1.0,
                           description="Some fluid cell covered by the moving body gets IB drag."),
            bounded_metric("ib_total_force_x", float(d["ib_total_force_x"]), upper=-1.0e-12,
                           description="The IB drag points against the +x flow."),
            bounded_metric("ib_total_force_norm", float(d["ib_total_force_norm"]), lower=1.0e-12,
                           description="The summed IB drag is not zero."),
            bounded_metric("max_velocity_not_increased", float(sim.lbm.max_velocity_norm() - speed0),
                           upper=1.0e-8, description="IB drag does not raise the largest fluid speed."),
            _finite_flag("velocity_finite", sim.lbm.velocity_numpy(),
                         "With IB drag on, no lattice velocity is NaN or infinite."),
        ),
        metadata={"drag": 0.1, "particle_volume": 4.0})


def run_contact_diagnostics_case() -> ValidationReport:
    # fsi/validation_cases.py:508-581: a particle inside a static 3^3 solid block
    cfg = SimulationConfig(
        num_steps=1, lbm_dt=1.0e-3, lbm=_periodic_lbm_config(initial_velocity=(0.0, 0.0, 0.0)),
        mpm=_small_mpm_config(),
        coupling=CouplingConfig(gamma=0.0, mpm_substeps_per_lbm_step=1, contact_enabled=True,
                                contact_velocity_damping=0.25, contact_fraction_threshold=0.01))
    solid = np.zeros((8, 8, 8), dtype=np.int32)
    solid[3:6, 3:6, 3:6] = 1
    sim = FSISimulation(cfg)
    sim.initialize_lbm(solid_np=solid)
    _one_particle(sim, velocity=(0.1, 0.0, 0.0))
    speed0 = sim.mpm.max_velocity_norm()
    sim.step()
    speed1 = sim.mpm.max_velocity_norm()
    d = sim.coupler.coupling_diagnostics()
    return ValidationReport(
        case_name="contact_diagnostics",
        metrics=(
            bounded_metric("contact_candidate_count", float(d["contact_candidate_count"]), lower=1.0,
                           description="A particle near the fixed solid is flagged for contact."),
            bounded_metric("contact_damped_particle_count", float(d["contact_damped_particle_count"]), lower=1.0,
                           description="The contact damping counter is non-zero."),
            bounded_metric("particle_contact_mask_sum", float(sim.coupler.particle_contact_mask_numpy().sum()),
                           lower=1.0, description="The contact mask of the active particle is 1."),
            bounded_metric("particle_speed_delta", float(speed1 - speed0), upper=-1.0e-6,
                           description="Contact damping slows the particle down."),
        ),
        metadata={"damping": 0.25, "initial_speed": speed0})


def run_validation_suite() -> list[ValidationReport]:
    # fsi/validation_cases.py:584-596, same order
    return [f() for f in (
        run_lbm_mass_conservation_case, run_lbm_force_response_case, run_mpm_zero_gravity_com_case,
        run_mpm_gravity_response_case, run_coupled_drift_case, run_coupling_force_balance_case,
        run_coupling_force_limit_case, run_coupling_boundary_support_case, run_immersed_boundary_drag_case,
        run_contact_diagnostics_case)]
