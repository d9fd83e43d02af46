"""Host-side logic that runs without a GPU: configuration validation (the
reference's tests/test_config.py rules), solver argument checks that raise
before any device work, the import surface, and the C ABI symbol table."""

import re
from pathlib import Path

import numpy as np
import pytest

from fsi.config import (
    BoundaryConfig,
    CouplingConfig,
    LBMConfig,
    MPMConfig,
    OutputConfig,
    SimulationConfig,
)

ROOT = Path(__file__).resolve().parents[1]


def test_defaults_match_reference():
    # fsi/config.py:20-259 defaults
    lbm = LBMConfig()
    assert (lbm.nx, lbm.ny, lbm.nz, lbm.viscosity, lbm.rho0) == (64, 32, 16, 0.1, 1.0)
    assert lbm.x_left.boundary_type == "velocity" and lbm.x_right.boundary_type == "pressure"
    mpm = MPMConfig()
    assert (mpm.max_particles, mpm.particles_per_cell, mpm.dx, mpm.dt) == (200_000, 2, 1.0, 0.25)
    assert mpm.gravity == (0.0, -9.8e-4, 0.0) and mpm.boundary_width == 3
    c = CouplingConfig()
    assert (c.gamma, c.mpm_substeps_per_lbm_step, c.min_valid_weight) == (0.05, 4, 1.0e-6)
    s = SimulationConfig(mpm=MPMConfig(nx=64, ny=32, nz=16))
    s.validate()
    assert s.mpm_dt == pytest.approx(0.25)


@pytest.mark.parametrize(
    "cfg",
    [
        BoundaryConfig("bogus"),
        BoundaryConfig(rho=0.0),
        LBMConfig(nx=4),
        LBMConfig(viscosity=0.0),
        LBMConfig(rho0=-1.0),
        LBMConfig(force=(1.0, 2.0)),
        LBMConfig(storage="int8"),
        MPMConfig(max_particles=0),
        MPMConfig(dx=0.0),
        MPMConfig(poisson_ratio=0.5),
        MPMConfig(material="plastic"),
        MPMConfig(boundary_width=0),
        MPMConfig(boundary_damping=1.5),
        CouplingConfig(gamma=-1.0),
        CouplingConfig(mpm_substeps_per_lbm_step=0),
        CouplingConfig(kernel="cubic"),
        CouplingConfig(force_limit=0.0),
        CouplingConfig(relative_velocity_limit=-1.0),
        CouplingConfig(gamma_ramp_steps=-1),
        CouplingConfig(min_valid_weight=0.0),
        CouplingConfig(min_valid_weight=1.5),
        CouplingConfig(immersed_boundary_drag=-1.0),
        CouplingConfig(immersed_boundary_fraction_threshold=1.1),
        CouplingConfig(immersed_boundary_max_force=0.0),
        CouplingConfig(contact_velocity_damping=-0.1),
        CouplingConfig(contact_fraction_threshold=1.1),
        OutputConfig(output_interval=0),
        OutputConfig(output_format="hdf5"),
        SimulationConfig(backend="tpu"),
        SimulationConfig(num_steps=0),
        SimulationConfig(lbm_dt=0.0),
        SimulationConfig(mpm=MPMConfig(nx=65)),
        CouplingConfig(mpm_refinement=0),
        CouplingConfig(mpm_refinement=1.5),
        # refined MPM grid: every axis r times the lattice's
        SimulationConfig(coupling=CouplingConfig(mpm_refinement=2)),
        SimulationConfig(coupling=CouplingConfig(mpm_refinement=2),
                         mpm=MPMConfig(nx=128, ny=128, nz=64)),
    ],
)
def test_invalid_config_raises_value_error(cfg):
    with pytest.raises(ValueError):
        cfg.validate()


def test_unsupported_lattice_features_raise_before_device_work():
    from fsi.lbm3d import LBMSolver3D

    periodic = {f: BoundaryConfig("periodic") for f in ("x_left", "x_right", "y_left", "y_right", "z_left", "z_right")}
    with pytest.raises(NotImplementedError):
        LBMSolver3D(LBMConfig(use_sparse=True, **periodic))
    with pytest.raises(NotImplementedError):
        LBMSolver3D(LBMConfig(**{**periodic, "y_left": BoundaryConfig("pressure")}))
    with pytest.raises(NotImplementedError):
        LBMSolver3D(LBMConfig(**{**periodic, "x_left": BoundaryConfig("wall")}))
    # 32-bit byte offsets into a population buffer: 384^3 FP32 is one cell
    # past the limit, FP16 halves the bytes (raised before any allocation)
    with pytest.raises(NotImplementedError, match="4 GiB"):
        LBMSolver3D(LBMConfig(nx=384, ny=384, nz=384, **periodic))
    with pytest.raises(NotImplementedError, match="4 GiB"):
        LBMSolver3D(LBMConfig(nx=512, ny=512, nz=512, storage="fp16", **periodic))


def test_solvers_fail_loudly_without_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from fsi.lbm3d import LBMSolver3D
    from fsi.mpm3d import MPMSolver3D

    periodic = {f: BoundaryConfig("periodic") for f in ("x_left", "x_right", "y_left", "y_right", "z_left", "z_right")}
    with pytest.raises(RuntimeError):
        LBMSolver3D(LBMConfig(nx=8, ny=8, nz=8, **periodic))
    with pytest.raises(RuntimeError):
        MPMSolver3D(MPMConfig(nx=8, ny=8, nz=8))


def test_import_surface():
    import fsi

    for name in ("SimulationConfig", "LBMConfig", "MPMConfig", "CouplingConfig", "BoundaryConfig",
                 "OutputConfig"):
        assert hasattr(fsi, name)
    for name in ("LBMSolver3D", "MPMSolver3D", "LBMMpmCoupler", "FSISimulation",
                 "SimulationOutputWriter"):
        assert name in fsi.__all__


def test_placeholder_modules_import(tmp_path):
    # fsi/io.py and fsi/units.py of the reference: a folder helper and an
    # empty unit class, kept importable for callers that name them
    from fsi.io import ensure_output_dir
    from fsi.units import LatticeUnits

    target = tmp_path / "a" / "b"
    ensure_output_dir(target)
    ensure_output_dir(target)
    assert target.is_dir()
    LatticeUnits()


def _declared_symbols():
    text = (ROOT / "include" / "fsi_hip.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(fsi_\w+)\s*\(", text, re.M)))


def test_c_abi_library_exports_every_declared_symbol():
    """The built library loads on the CPU and exports each entry point of the header."""
    from fsi import _native

    lib = _native.load_library()
    names = _declared_symbols()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name
    assert lib.fsi_abi_version() == 5
    assert set(names) >= set(_native._SIGNATURES)


def test_lbm_load_solid_from_txt(tmp_path):
    # fsi/lbm3d.py load_solid_from_txt: whitespace text, Fortran-order reshape
    from fsi.lbm3d import LBMSolver3D

    solid = (np.arange(4 * 3 * 2) % 3 == 0).astype(np.int32).reshape(4, 3, 2)
    path = tmp_path / "mask.txt"
    np.savetxt(path, solid.reshape(-1, order="F"), fmt="%d")
    got = LBMSolver3D.load_solid_from_txt(str(path), 4, 3, 2)
    assert np.array_equal(got, solid)


def test_bench_command_line_on_the_cpu():
    """bench.py parses its arguments without a GPU and refuses more than one
    GPU with a message (this build runs on one)."""
    import subprocess
    import sys
    from pathlib import Path

    bench = Path(__file__).resolve().parents[1] / "bench.py"
    r = subprocess.run([sys.executable, str(bench), "--help"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "--large-error-steps" in r.stdout
    r = subprocess.run([sys.executable, str(bench), "--gpus", "8"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_bench_scenes_as_the_configurations_state_them():
    """The coupled scenes bench.py measures: the 32^3 cube and the 16-sphere
    suspension at dx_mpm = dx_lbm / 2 (8 particles per MPM cell), their
    co-located counterparts, and the elastic P-wave Courant number 0.3 that
    sets their substep counts."""
    import numpy as np

    import bench
    from fsi.scenes import body_moduli, sphere_pack

    def courant(cfg, E):
        nu, rho = cfg.mpm.poisson_ratio, cfg.mpm.density
        c_p = np.sqrt(E * (1 - nu) / ((1 + nu) * (1 - 2 * nu)) / rho)
        return c_p * (cfg.lbm_dt / cfg.coupling.mpm_substeps_per_lbm_step) / cfg.mpm.dx

    for refine, lo, hi, sub in ((2, (10, 14, 14), (14, 18, 18), 224), (1, (8, 12, 12), (16, 20, 20), 112)):
        cfg, solid, pos = bench.coupled_scene_32(refine=refine)
        cfg.validate()
        assert len(pos) == 4096 and cfg.coupling.mpm_refinement == refine
        assert (cfg.mpm.nx, cfg.mpm.dx) == (32 * refine, 1.0 / refine)
        assert cfg.coupling.mpm_substeps_per_lbm_step == sub
        assert np.all(pos >= np.array(lo)) and np.all(pos <= np.array(hi))
        assert np.allclose(pos.mean(axis=0), (12.0, 16.0, 16.0), atol=0.05)
        assert bench.particle_volume(cfg) == (0.5 / refine) ** 3
        assert 0.25 < courant(cfg, cfg.mpm.youngs_modulus) <= 0.3

    fine = bench.coupled_scene_large(refine=2)
    flat = bench.coupled_scene_large(refine=1)
    fine[0].validate()
    assert len(fine[2]) == 3591333 and len(flat[2]) == 448861
    assert (fine[0].mpm.nx, fine[0].mpm.ny, fine[0].mpm.nz) == (512, 256, 256)
    # the same 16 spheres, 8x the particles, the same per-sphere moduli
    _, info2 = sphere_pack((256, 128, 128), 16, 6.0, 12.0, spacing=0.25, seed=0)
    _, info1 = sphere_pack((256, 128, 128), 16, 6.0, 12.0, spacing=0.5, seed=0)
    assert np.array_equal(info1["centres"], info2["centres"]) and np.array_equal(info1["radii"], info2["radii"])
    assert np.array_equal(body_moduli(info2["counts"], 5.0e2, 5.0e3, seed=0)[1],
                          body_moduli(info1["counts"], 5.0e2, 5.0e3, seed=0)[1])
    assert np.isclose(fine[3].max(), flat[3].max())
    for scene in (fine, flat):
        assert 0.25 < courant(scene[0], scene[3].max()) <= 0.3
