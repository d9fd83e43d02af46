"""Scenario runners and field comparisons shared by the parity tests.

A scenario builds the same seeded scene through an API namespace (oracle or
product, see cases.py), runs it, and returns the reference-layout fields as
numpy arrays, so oracle and HIP results compare entry by entry.
"""

from __future__ import annotations

import numpy as np

from fsi.config import BoundaryConfig, CouplingConfig, LBMConfig, MPMConfig, SimulationConfig

W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)


def rel_inf(a, b) -> float:
    """max |a - b| / max |b| (the per-field norm of the parity statements)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = float(np.abs(b).max()) if b.size else 0.0
    if scale == 0.0:
        return float(np.abs(a - b).max()) if a.size else 0.0
    return float(np.abs(a - b).max() / scale)


def close(a, b, rel, abs_floor) -> tuple[bool, float, float]:
    """max|a - b| <= rel * max|b| + abs_floor; returns (ok, rel error, abs error)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.size == 0:
        return True, 0.0, 0.0
    diff = float(np.abs(a - b).max())
    scale = float(np.abs(b).max())
    return diff <= rel * scale + abs_floor, (diff / scale if scale else diff), diff


def lbm_fields(solver, rho0=1.0) -> dict:
    f = solver.distribution_numpy()
    return {
        "dev": f - W * rho0,
        "rho": solver.density_numpy(),
        "u": solver.velocity_numpy(),
    }


def channel_config(nx=32, ny=16, nz=16, storage=None, **kw) -> tuple[LBMConfig, np.ndarray]:
    """Inflow/outflow channel: velocity x-, pressure x+, walls as a solid shell in y/z,
    an obstacle block, as in the coupled scenes of the benchmark configs."""
    params = dict(
        nx=nx, ny=ny, nz=nz, viscosity=0.1 / 3, rho0=1.0,
        x_left=BoundaryConfig("velocity", velocity=(0.02, 0.0, 0.0)),
        x_right=BoundaryConfig("pressure", rho=1.0),
    )
    params.update(kw)
    if storage is not None:
        params["storage"] = storage
    cfg = LBMConfig(**params)
    solid = np.zeros((nx, ny, nz), np.int32)
    solid[:, 0, :] = 1
    solid[:, -1, :] = 1
    solid[:, :, 0] = 1
    solid[:, :, -1] = 1
    solid[nx // 3: nx // 3 + 4, ny // 2 - 2: ny // 2 + 2, nz // 2 - 2: nz // 2 + 2] = 1
    return cfg, solid


def porous_mask(n, seed=0, porosity=0.8, rmin=2.0, rmax=4.0):
    """Random periodic sphere packing voxelised at cell centres (solid fraction >= 1 - porosity)."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = n
    solid = np.zeros(n, bool)
    grid = np.stack(np.meshgrid(np.arange(nx) + 0.5, np.arange(ny) + 0.5, np.arange(nz) + 0.5,
                                indexing="ij"), axis=-1)
    dims = np.array(n, dtype=np.float64)
    while solid.mean() < 1.0 - porosity:
        c = rng.uniform(0, 1, 3) * dims
        r = rng.uniform(rmin, rmax)
        d = grid - c
        d -= dims * np.round(d / dims)
        solid |= (d ** 2).sum(-1) <= r * r
    return solid.astype(np.int32)


def run_lbm(api_lbm, cfg, steps, solid=None, force_field=None, force_change=None):
    s = api_lbm(cfg)
    s.initialize(solid)
    if force_field is not None:
        s.set_force_from_numpy(force_field)
    for i in range(steps):
        if force_change is not None and i == steps // 2:
            force_change(s)
        s.step()
    return s


def jittered_cube(lower, upper, spacing, seed, jitter=0.25):
    """Lattice particles in [lower, upper) with U(-jitter/2, jitter/2)*spacing offsets."""
    rng = np.random.default_rng(seed)
    axes = [np.arange(lower[a], upper[a], spacing) + 0.5 * spacing for a in range(3)]
    mesh = np.meshgrid(*axes, indexing="ij")
    pos = np.stack([m.ravel() for m in mesh], axis=1)
    pos += rng.uniform(-0.5 * jitter, 0.5 * jitter, pos.shape) * spacing
    return pos.astype(np.float32)


def coupled_scene(storage=None, n=(24, 16, 16), substeps=8, gamma=0.5, e_mod=400.0,
                  ib=False, contact=False, seed=0):
    """Small coupled penalty scene: channel flow past an elastic jittered cube."""
    nx, ny, nz = n
    lbm_cfg, solid = channel_config(nx, ny, nz, storage=storage, initial_velocity=(0.02, 0.0, 0.0))
    # no obstacle block in the coupled scene: the solid is the MPM cube
    solid[:, 1:-1, 1:-1] = 0
    mpm_cfg = MPMConfig(nx=nx, ny=ny, nz=nz, max_particles=4096, dx=1.0, dt=0.1, density=1.2,
                        youngs_modulus=e_mod, poisson_ratio=0.3, gravity=(0.0, 0.0, 0.0),
                        boundary_width=2)
    cpl = CouplingConfig(gamma=gamma, mpm_substeps_per_lbm_step=substeps,
                         immersed_boundary_enabled=ib, immersed_boundary_drag=0.05 if ib else 0.0,
                         immersed_boundary_fraction_threshold=0.05,
                         contact_enabled=contact, contact_velocity_damping=0.1 if contact else 0.0,
                         contact_fraction_threshold=0.9)
    sim_cfg = SimulationConfig(num_steps=10, lbm_dt=0.1, lbm=lbm_cfg, mpm=mpm_cfg, coupling=cpl)
    c = np.array([nx * 0.35, ny / 2, nz / 2])
    pos = jittered_cube(c - 3, c + 3, 0.5, seed)
    omega = np.array([0.0, 0.01, 0.02])
    vel = np.cross(omega, pos - c).astype(np.float32)
    return sim_cfg, solid, pos, vel


def refined_scene(storage=None, refine=2, n=(16, 12, 12), substeps=16, ib=False, contact=False,
                  seed=0, wall=False):
    """The coupled scene with the MPM grid ``refine`` times finer than the
    lattice (MPM dx = 1/refine lattice cells, 2 particles per MPM cell and
    axis): a jittered cube of side 5 lattice cells in the channel, spinning;
    ``wall`` puts it against the solid shell at y = 0 (partial support).
    ``refine=1`` gives the same particles on a co-located MPM grid."""
    nx, ny, nz = n
    lbm_cfg, solid = channel_config(nx, ny, nz, storage=storage, initial_velocity=(0.02, 0.0, 0.0))
    solid[:, 1:-1, 1:-1] = 0
    dx = 1.0 / refine
    mpm_cfg = MPMConfig(nx=refine * nx, ny=refine * ny, nz=refine * nz, max_particles=8192, dx=dx,
                        dt=0.1, density=1.2, youngs_modulus=400.0, poisson_ratio=0.3,
                        gravity=(0.0, 0.0, 0.0), boundary_width=2)
    cpl = CouplingConfig(gamma=0.5, mpm_substeps_per_lbm_step=substeps, mpm_refinement=refine,
                         immersed_boundary_enabled=ib, immersed_boundary_drag=0.05 if ib else 0.0,
                         immersed_boundary_fraction_threshold=0.05,
                         contact_enabled=contact, contact_velocity_damping=0.1 if contact else 0.0,
                         contact_fraction_threshold=0.9)
    sim_cfg = SimulationConfig(num_steps=10, lbm_dt=0.1, lbm=lbm_cfg, mpm=mpm_cfg, coupling=cpl)
    c = np.array([nx * 0.35, 3.8 if wall else ny / 2, nz / 2])
    pos = jittered_cube(c - 2.5, c + 2.5, 0.25, seed)
    omega = np.array([0.0, 0.01, 0.02])
    vel = np.cross(omega, pos - c).astype(np.float32)
    return sim_cfg, solid, pos, vel


REFINED_MASS, REFINED_VOLUME = 1.2 * 0.25 ** 3, 0.25 ** 3


def sim_fields(sim) -> dict:
    out = lbm_fields(sim.lbm)
    out.update(
        x=sim.mpm.positions_numpy(),
        v=sim.mpm.velocities_numpy(),
        F=sim.mpm.deformation_gradients_numpy(),
        pforce=sim.mpm.particle_forces_numpy(),
        cforce=sim.coupler.coupling_force_numpy(),
        phi=sim.coupler.solid_volume_fraction_numpy(),
        lbm_force=sim.lbm.force_numpy(),
        cmask=sim.coupler.particle_coupling_mask_numpy(),
        wvalid=sim.coupler.particle_valid_weights_numpy(),
    )
    return out
