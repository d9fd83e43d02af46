#!/usr/bin/env python
"""Benchmark of the coupled LBM-MPM step on one MI355X.

Workload (the lattice bandwidth case): D3Q19 MRT on a 256^3 lattice with a
random porous solid mask (porosity 0.8, spheres of radius 4-10 cells, seed 0),
all faces periodic, body force g_x = 1e-6, tau = 0.55. A step is one lattice
update of the whole 256^3 lattice (one fused collide+stream launch); `value`
is the FP32-population MLUPS (all cells), the reference's arithmetic type.
The same scene then runs with FP16 shifted storage and with I16 fixed-point
storage (untimed 1,000-step accuracy runs against the FP64 build), the 32^3
coupled penalty scene (4096 particles, 112 MPM substeps per lattice step)
and the 256x128x128 suspension (16 spheres of per-sphere E, 256 substeps)
report coupled steps/s, the suspension also against its FP64 build. Inputs
are resident in HBM before the timed region.

Prints ONE JSON line. `--gpus` above 1 is refused: this build runs on one GPU.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for _p in (ROOT, ROOT / "taichi-mpm-lbm_amd"):
    if str(_p) not in sys.path:
        sys.path.insert(0, str(_p))

METRIC = "LBM MLUPS (%HBM roofline) + coupled steps/s @256³; FP16-store/learned vs FP32"
HBM_PEAK_GBS = 8000.0


def periodic_lbm_config(n, storage, tau=0.55, force=(1.0e-6, 0.0, 0.0)):
    from fsi.config import BoundaryConfig, LBMConfig

    faces = {f: BoundaryConfig("periodic") for f in ("x_left", "x_right", "y_left", "y_right", "z_left", "z_right")}
    return LBMConfig(nx=n, ny=n, nz=n, viscosity=(tau - 0.5) / 3.0, rho0=1.0, force=force,
                     storage=storage, **faces)


def algorithmic_bytes(n_fluid, n_cells, pop_bytes):
    # per step: read + write 19 populations per fluid cell, 1 mask byte per cell
    return n_fluid * 2 * 19 * pop_bytes + n_cells * 1


def warm_up(step, count, min_seconds=0.25):
    """`count` untimed steps, continued until `min_seconds` have passed (the
    clocks and caches reach their steady state in a fraction of a second; a
    short count alone under-reports the fast storages)."""
    import torch

    t0 = time.perf_counter()
    done = 0
    while done < count or time.perf_counter() - t0 < min_seconds:
        step()
        done += 1
        if done % 20 == 0:
            torch.cuda.synchronize()
    torch.cuda.synchronize()
    return done


def timed_loop(step, steps):
    """Wall time of exactly `steps` steps between two synchronisations, and
    the same interval from HIP events on the launch stream."""
    import torch

    stream = torch.cuda.current_stream()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    a.record(stream)
    for _ in range(steps):
        step()
    b.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall, a.elapsed_time(b) * 1e-3


def time_lattice(solver, steps, warmup, launch_samples):
    import torch

    warm_up(solver.step, warmup)
    wall, _ = timed_loop(solver.step, steps)
    # per-launch kernel duration: event pairs on the stream the kernel runs on
    stream = torch.cuda.current_stream()
    pairs = []
    for _ in range(launch_samples):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        solver.step()
        b.record(stream)
        pairs.append((a, b))
    torch.cuda.synchronize()
    launch_ms = sum(a.elapsed_time(b) for a, b in pairs) / len(pairs)
    return wall, launch_ms


def lattice_errors_vs_fp64(n, mask, fields, steps):
    """Per-field error of the timed lattice runs (continued, untimed, to
    `steps` steps) against the FP64 build of the same kernels after the same
    number of steps (the FP64 build is pinned to the CPU oracle at 1e-10):
    max-norm relative error of u, its L2-relative error, and the max-norm
    error of rho (absolute and relative to max |rho - rho0| of the FP64 run),
    computed on the device."""
    import torch

    import fsi

    s = fsi.LBMSolver3D(periodic_lbm_config(n, "fp64"))
    s.initialize(mask)
    for _ in range(steps):
        s.step()
    s._ensure_macro()
    rho64, vel64 = s.rho, s.vel
    drho64 = (rho64 - 1.0).abs().max().item()
    out = {"steps": steps, "reference": "FP64 build of the same kernels, same mask and steps"}
    out["u_max"] = float(vel64.abs().max().item())
    for storage, (rho, vel, extra) in fields.items():
        dv = vel.double() - vel64
        out[storage] = {
            "u_rel_max": float(dv.abs().max().item() / vel64.abs().max().item()),
            "u_rel_l2": float(dv.norm().item() / vel64.norm().item()),
            "rho_abs_max": float((rho.double() - rho64).abs().max().item()),
            "rho_rel_to_deviation": float((rho.double() - rho64).abs().max().item() / max(drho64, 1e-300)),
            **extra,
        }
    del s
    torch.cuda.empty_cache()
    return out


def read_traffic(name):
    """HBM bytes per launch from the committed rocprofv3 PMC summary, if present."""
    path = ROOT / "profiles" / name
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def _cpu_task(task, budget_s):
    """One CPU measurement of the oracle (C++ restatement of the reference's
    algorithm in its operation order, FP32), run in a child process so its
    OpenMP thread count is its own. Prints one JSON object."""
    import numpy as np

    from fsi.scenes import porous_spheres
    from oracle import pyoracle as po

    po.build()
    out = {"threads": int(os.environ.get("OMP_NUM_THREADS", len(os.sched_getaffinity(0))))}
    if task == "lattice":
        mask, _ = porous_spheres((256, 256, 256), porosity=0.8, rmin=4.0, rmax=10.0, seed=0)
        s = po.OracleLBM(periodic_lbm_config(256, "fp32"), "f32")
        s.initialize(mask)
        s.step()  # warm-up (page faults, thread start)
        done, t0 = 0, time.perf_counter()
        while done < 40 and (done == 0 or time.perf_counter() - t0 < budget_s):
            s.step()
            done += 1
        dt = time.perf_counter() - t0
        out.update(value=mask.size * done / dt / 1e6, unit="MLUPS", steps=done, seconds=dt)
    elif task == "coupled_32":
        cfg, solid, pos = coupled_scene_32()
        sim = po.OracleSimulation(cfg, "f32")
        sim.initialize_lbm(solid)
        sim.initialize_mpm_from_numpy(pos, particle_volume=particle_volume(cfg))
        sim.step()
        done, t0 = 0, time.perf_counter()
        while done < 20 and (done == 0 or time.perf_counter() - t0 < budget_s):
            sim.step()
            done += 1
        dt = time.perf_counter() - t0
        out.update(value=done / dt, unit="coupled steps/s", steps=done, seconds=dt)
    elif task == "coupled_large":
        # the suspension's coupled step: one lattice step plus 256 MPM
        # substeps with the coupling; timed as one lattice step and two
        # coupled substeps, extrapolated to 256 substeps (the CPU oracle has
        # the reference's one global material; the cost does not depend on E)
        import dataclasses

        cfg, solid, pos, _ = coupled_scene_large(refine=LARGE_REFINE)
        sub = cfg.coupling.mpm_substeps_per_lbm_step
        cfg2 = dataclasses.replace(cfg, coupling=dataclasses.replace(cfg.coupling, mpm_substeps_per_lbm_step=2))
        sim = po.OracleSimulation(cfg2, "f32")
        sim.initialize_lbm(solid)
        sim.initialize_mpm_from_numpy(pos, particle_volume=particle_volume(cfg))
        t0 = time.perf_counter()
        sim.lbm.step()
        t_lat = time.perf_counter() - t0
        t0 = time.perf_counter()
        sim.coupler.step(cfg.lbm_dt)  # 2 substeps + force assembly + a lattice step
        t_two = time.perf_counter() - t0 - t_lat
        per_sub = t_two / 2.0
        per_step = t_lat + sub * per_sub
        out.update(value=1.0 / per_step, unit="coupled steps/s", lattice_step_s=t_lat,
                   substep_s=per_sub, extrapolated_step_s=per_step, particles=int(len(pos)),
                   sample=f"1 lattice step + 1 coupled step of 2 substeps, extrapolated to {sub} substeps")
    print(json.dumps(out), flush=True)


def cpu_baseline(budget_s=10.0):
    """The CPU oracle on the host cores beside the GPU run, in child processes:
    the 256^3 lattice on all cores (the headline comparison) and on one core,
    the 32^3 coupled scene on all cores and on one core, and the 256x128x128
    suspension's coupled step on all cores (extrapolated from a sample)."""
    import subprocess

    cores = len(os.sched_getaffinity(0))

    def run(task, threads):
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        r = subprocess.run([sys.executable, str(Path(__file__).resolve()), "--cpu-task", task,
                            "--cpu-budget", str(budget_s)], env=env, capture_output=True, text=True,
                           timeout=600)
        if r.returncode != 0:
            return {"error": r.stderr[-400:]}
        return json.loads(r.stdout.strip().splitlines()[-1])

    lat = run("lattice", cores)
    lat1 = run("lattice", 1)
    c32 = run("coupled_32", cores)
    c32_1 = run("coupled_32", 1)
    large = run("coupled_large", cores)
    main = {
        "value": round(lat.get("value", float("nan")), 3),
        "unit": "MLUPS",
        "cores": int(lat.get("threads", cores)),
        "kind": "port",
        "sample": f"{lat.get('steps')} timed steps (+1 warm-up) of the 256^3 porous lattice in the CPU oracle "
                  f"(C++ restatement in the reference's operation order), FP32, OpenMP; "
                  f"{lat.get('seconds', 0):.1f} s",
    }
    more = {
        "lattice_1_core": {"value": round(lat1.get("value", float("nan")), 3), "unit": "MLUPS", "cores": 1,
                           "sample": f"{lat1.get('steps')} timed steps of the same lattice"},
        "coupled_32": {"value": round(c32.get("value", float("nan")), 4), "unit": "coupled steps/s",
                       "cores": int(c32.get("threads", cores)),
                       "sample": f"{c32.get('steps')} coupled steps of the 32^3 scene (4096 particles, MPM grid "
                                 f"64^3 at dx_mpm = dx_lbm / 2, 224 substeps)"},
        "coupled_32_1_core": {"value": round(c32_1.get("value", float("nan")), 4), "unit": "coupled steps/s",
                              "cores": 1, "sample": f"{c32_1.get('steps')} coupled steps"},
        "coupled_large": {"value": round(large.get("value", float("nan")), 5), "unit": "coupled steps/s",
                          "cores": int(large.get("threads", cores)),
                          "sample": large.get("sample"), "extrapolated_step_s": large.get("extrapolated_step_s"),
                          "particles": large.get("particles")},
    }
    return main, more


def coupled_scene_32(storage="fp32", refine=2):
    """32^3 coupled penalty scene: inflow u_x=0.02 (velocity face x-), pressure
    face x+, no-slip walls as a solid shell in y/z, tau=0.6; an elastic cube
    of 8^3 MPM cells with 8 particles per cell (16^3 = 4096 particles,
    spacing dx_mpm / 2, jitter U(-1/4, 1/4) of the spacing, seed 0) centred
    at (12, 16, 16), E=1e3, nu=0.3, density 1.2, gamma=0.5. refine=2: the MPM
    grid is 64^3 at dx_mpm = dx_lbm / 2 (the cube spans [10, 14] x [14, 18]^2,
    224 substeps); refine=1: co-located, the cube spans [8, 16] x [12, 20]^2,
    112 substeps (P-wave Courant number 0.3 either way)."""
    import numpy as np

    from fsi.config import BoundaryConfig, CouplingConfig, LBMConfig, MPMConfig, SimulationConfig
    from fsi.scenes import channel_walls, jittered_box

    n = 32
    dx = 1.0 / refine
    sub = 112 * refine
    lbm = LBMConfig(nx=n, ny=n, nz=n, viscosity=0.1 / 3.0, rho0=1.0,
                    x_left=BoundaryConfig("velocity", velocity=(0.02, 0.0, 0.0)),
                    x_right=BoundaryConfig("pressure", rho=1.0), storage=storage)
    mpm = MPMConfig(nx=refine * n, ny=refine * n, nz=refine * n, max_particles=4096, dx=dx, dt=1.0 / sub,
                    density=1.2, youngs_modulus=1.0e3, poisson_ratio=0.3, gravity=(0.0, 0.0, 0.0),
                    boundary_width=2)
    cpl = CouplingConfig(gamma=0.5, mpm_substeps_per_lbm_step=sub, mpm_refinement=refine)
    cfg = SimulationConfig(num_steps=100, lbm_dt=1.0, lbm=lbm, mpm=mpm, coupling=cpl)
    solid = channel_walls((n, n, n))
    h = 4.0 * dx  # half the cube's side: 8 MPM cells
    centre = np.array([12.0, 16.0, 16.0])
    pos = jittered_box(tuple(centre - h), tuple(centre + h), 0.5 * dx, seed=0, jitter=0.5)
    return cfg, solid, pos


def particle_volume(cfg):
    """Volume of one of the 8 particles per MPM cell of these scenes."""
    return (0.5 * cfg.mpm.dx) ** 3


LARGE_SUBSTEPS = 256
LARGE_REFINE = 2  # the suspension as stated: dx_mpm = dx_lbm / 2


def large_substeps(refine=1):
    """Substeps per lattice step of the suspension: the stiffest sphere's
    P-wave speed sqrt((lam + 2 mu)/rho) is 74.9 lattice cells per time unit,
    and mpm_dt <= 0.3 dx_mpm / c_p needs >= 250 substeps at dx_mpm = dx_lbm
    (256 here), twice that at dx_mpm = dx_lbm / 2."""
    return LARGE_SUBSTEPS * refine


def coupled_scene_large(storage="fp32", refine=1):
    """256x128x128 coupled suspension: inflow u_x=0.04 (velocity face x-,
    fluid starting at that velocity), pressure face x+, no-slip walls as a
    solid shell in y/z, tau=0.56; 16 non-overlapping elastic spheres r U[6,12]
    lattice cells (sphere_pack seed 0), 8 particles per MPM cell (spacing
    dx_mpm / 2, jitter U(-1/4, 1/4) of the spacing), E log-uniform in
    [5e2, 5e3] per sphere (seed 0, a per-particle material plane), nu=0.3,
    density 1.2, gamma=0.5. refine=2: the MPM grid is 512x256x256 at
    dx_mpm = dx_lbm / 2 (the configuration as stated), refine=1 co-located.
    Returns (config, solid mask, positions, per-particle E)."""
    from fsi.config import BoundaryConfig, CouplingConfig, LBMConfig, MPMConfig, SimulationConfig
    from fsi.scenes import body_moduli, channel_walls, sphere_pack

    shape = (256, 128, 128)
    dx = 1.0 / refine
    pos, info = sphere_pack(shape, 16, 6.0, 12.0, spacing=0.5 * dx, seed=0)
    E, _ = body_moduli(info["counts"], 5.0e2, 5.0e3, seed=0)
    # the fluid starts at the inflow velocity: started from rest, the
    # reference's velocity face on this channel becomes unstable after ~140
    # steps (reproduced by the CPU oracle in the reference's operation order)
    lbm = LBMConfig(nx=shape[0], ny=shape[1], nz=shape[2], viscosity=0.06 / 3.0, rho0=1.0,
                    initial_velocity=(0.04, 0.0, 0.0),
                    x_left=BoundaryConfig("velocity", velocity=(0.04, 0.0, 0.0)),
                    x_right=BoundaryConfig("pressure", rho=1.0), storage=storage)
    sub = large_substeps(refine)
    mpm = MPMConfig(nx=refine * shape[0], ny=refine * shape[1], nz=refine * shape[2], max_particles=len(pos),
                    dx=dx, dt=1.0 / sub, density=1.2, youngs_modulus=1.0e3, poisson_ratio=0.3,
                    gravity=(0.0, 0.0, 0.0), boundary_width=2)
    cpl = CouplingConfig(gamma=0.5, mpm_substeps_per_lbm_step=sub, mpm_refinement=refine)
    cfg = SimulationConfig(num_steps=2000, lbm_dt=1.0, lbm=lbm, mpm=mpm, coupling=cpl)
    return cfg, channel_walls(shape), pos, E


def large_workload(refine):
    n = "512x256x256 MPM grid at dx_mpm = dx_lbm / 2" if refine == 2 else "co-located 256x128x128 MPM grid"
    return (f"256x128x128 coupled suspension, 16 elastic spheres (E log-uniform [5e2, 5e3] per sphere), "
            f"{n}, 8 particles per MPM cell, {large_substeps(refine)} substeps")


def large_sim(storage, refine=1):
    import fsi

    cfg, solid, pos, E = coupled_scene_large(storage, refine)
    sim = fsi.FSISimulation(cfg)
    sim.initialize_lbm(solid)
    sim.initialize_mpm_from_numpy(pos, particle_volume=(0.5 / refine) ** 3)
    sim.mpm.set_particle_material(E)
    return sim, cfg, pos


def time_coupled_large(steps, storage="fp32", min_warmup_s=0.5, refine=1):
    """Coupled steps/s of the suspension: warm-up of at least two steps and
    min_warmup_s seconds, then `steps` timed steps."""
    import torch

    sim, cfg, pos = large_sim(storage, refine)
    warm_up(sim.advance, 2, min_warmup_s)
    wall, _ = timed_loop(sim.advance, steps)
    n_sub = cfg.coupling.mpm_substeps_per_lbm_step
    return {"steps_per_s": round(steps / wall, 3), "ms_per_step": round(1e3 * wall / steps, 3),
            "particles": int(len(pos)), "substeps": n_sub, "timed_steps": steps,
            "particle_substeps_per_s": round(len(pos) * n_sub * steps / wall, 1)}


def large_errors_vs_fp64(steps, storages=("fp32", "i16"), refine=1):
    """The suspension after `steps` coupled steps from the same start in the
    FP32 build and the mixed-precision build (I16 lattice populations, FP32
    MPM and coupling), against the FP64 build: max-norm relative error of the
    lattice velocity and the particle velocity, the particle position error in
    cells and relative to the largest displacement, and aggregate measures
    that a chaotic suspension keeps longer: the mean streamwise velocity, the
    drag (the particles' step-averaged coupling force summed) and each
    sphere's centroid."""
    import numpy as np
    import torch

    from fsi.scenes import sphere_pack

    _, info = sphere_pack((256, 128, 128), 16, 6.0, 12.0, spacing=0.5 / refine, seed=0)
    bounds = np.concatenate([[0], np.cumsum(info["counts"])])

    def centroids(x):
        return np.stack([x[bounds[i]:bounds[i + 1]].mean(0) for i in range(len(bounds) - 1)])

    def run(storage):
        sim, _, pos = large_sim(storage, refine)
        sim.coupler.record_step_force()
        t0 = time.perf_counter()
        for i in range(steps):
            sim.advance()
            if (i + 1) % 100 == 0:  # progress of long runs
                torch.cuda.synchronize()
                print(f"[suspension {storage}] {i + 1}/{steps} steps, {time.perf_counter() - t0:.0f} s",
                      file=sys.stderr, flush=True)
        out = {"u": sim.lbm.velocity_numpy().astype(np.float64), "x": sim.mpm.positions_numpy().astype(np.float64),
               "v": sim.mpm.velocities_numpy().astype(np.float64), "x0": np.asarray(pos, np.float64),
               "drag": sim.coupler.step_force_numpy().astype(np.float64).sum(0)}
        fluid = sim.lbm.solid_numpy() == 0
        out["mean_ux"] = float(out["u"][..., 0][fluid].mean())
        out["centroids"] = centroids(out["x"])
        del sim
        torch.cuda.empty_cache()
        return out

    ref = run("fp64")
    disp = np.abs(ref["x"] - ref["x0"]).max()
    res = {"steps": steps, "reference": "FP64 build of the same kernels, same scene and steps",
           "u_max": float(np.abs(ref["u"]).max()), "particle_displacement_max_cells": float(disp)}
    for storage in storages:
        got = run(storage)
        res[storage] = {
            "u_rel_max": float(np.abs(got["u"] - ref["u"]).max() / max(np.abs(ref["u"]).max(), 1e-30)),
            "particle_v_rel_max": float(np.abs(got["v"] - ref["v"]).max() / max(np.abs(ref["v"]).max(), 1e-30)),
            "particle_x_abs_max_cells": float(np.abs(got["x"] - ref["x"]).max()),
            "particle_x_rel_to_max_displacement": float(np.abs(got["x"] - ref["x"]).max() / max(disp, 1e-30)),
            "mean_ux_rel": float(abs(got["mean_ux"] - ref["mean_ux"]) / abs(ref["mean_ux"])),
            "drag_x_rel": float(abs(got["drag"][0] - ref["drag"][0]) / abs(ref["drag"][0])),
            "sphere_centroid_abs_max_cells": float(np.abs(got["centroids"] - ref["centroids"]).max()),
        }
    res["sphere_centroid_displacement_max_cells"] = float(np.abs(ref["centroids"] - centroids(ref["x0"])).max())
    return res


def time_closure(train_scenes=8, train_steps=40, eval_steps=60, large_eval_steps=40, speed_steps=2):
    """Learned closure: trains on coarse/2x-refined pairs (single-ellipsoid
    channels, 48x24x24 coarse), then on a held-out ellipsoid
    scene and on the 64-sphere channel (192x96x96 coarse, FP16 populations;
    384x192x192 refined, FP32) compares drag and centroid displacement of
    the plain coarse run and of the coarse run with the closure against the
    refined run, and times the coarse run with the closure against the
    refined run."""
    import numpy as np
    import torch

    from fsi.closure import ClosureScene, PackScene, accuracy_vs_refined, scene_pair_data, train_closure

    t0 = time.perf_counter()
    # adding four pairs of four-sphere scenes made the held-out ellipsoid run
    # worse (drag error 0.73% instead of 0.54%, particle force error 14.0%
    # instead of 13.0%) and the 64-sphere run no better: ellipsoids only
    scenes = [ClosureScene.random(s, shape=(48, 24, 24), axes=(3.0, 5.0)) for s in range(300, 300 + train_scenes)]
    data = [scene_pair_data(sc, train_steps, every=2, warmup=6) for sc in scenes]
    X = np.concatenate([d[0] for d in data])
    T = np.concatenate([d[1] for d in data])
    t1 = time.perf_counter()
    closure, hist = train_closure(X, T, steps=1500, batch=16384, lr=2e-3, seed=1)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    held_out = accuracy_vs_refined(ClosureScene.random(400, shape=(48, 24, 24), axes=(3.0, 5.0)), eval_steps,
                                   closure)
    large = PackScene(shape=(192, 96, 96), count=64, radii=(2.5, 7.0), youngs_modulus=1.0e3, poisson_ratio=0.3,
                      inflow=0.04, margin=4.0, gap=2.0, seed=5)
    large_acc = accuracy_vs_refined(large, large_eval_steps, closure, storage="fp16")
    coarse = large.simulation(False, "fp16")
    coarse.coupler.set_closure(closure)
    fine_sim = large.simulation(True)

    def timed(sim, n):
        sim.step()
        torch.cuda.synchronize()
        a = time.perf_counter()
        for _ in range(n):
            sim.step()
        torch.cuda.synchronize()
        return (time.perf_counter() - a) / n

    tc = timed(coarse, speed_steps)
    tf = timed(fine_sim, 2 * speed_steps) * 2.0
    return {"train": {"scenes": len(scenes), "samples": int(len(X)), "adam_steps": 1500,
                      "data_s": round(t1 - t0, 2), "train_s": round(t2 - t1, 2),
                      "loss_first": round(hist[0], 4), "loss_last": round(hist[-1], 4)},
            "accuracy_vs_refined": {
                "metric": "drag = mean fluid-on-solid x-force per step (substep-averaged coupling force); "
                          "displacement = per-body centroid displacement vector error",
                "ellipsoid": {"scene": f"ellipsoid seed 400, 48x24x24 coarse FP32, {eval_steps} coarse steps",
                              **held_out},
                "spheres64": {"scene": f"64 spheres, 192x96x96 coarse FP16 vs 384x192x192 FP32, "
                                       f"{large_eval_steps} coarse steps", **large_acc}},
            "speed": {"workload": "64 spheres r 2.5-7 (coarse cells), 192x96x96 FP16 lattice + closure vs "
                                  "384x192x192 FP32 refined, 112 substeps",
                      "coarse_particles": int(coarse.mpm.num_particles),
                      "fine_particles": int(fine_sim.mpm.num_particles),
                      "coarse_closure_ms_per_step": round(1e3 * tc, 3),
                      "fine_ms_per_coarse_step": round(1e3 * tf, 3), "speedup": round(tf / tc, 2),
                      "closure_fallbacks": coarse.coupler.closure_fallback_count()}}


CFG4_SCENES = dict(shape=(96, 48, 48), axes=(4.0, 10.0), modulus=(2.0e2, 2.0e4), poisson=(0.2, 0.4),
                   inflow=(0.01, 0.05), substeps="stable")


def closure_data_and_training(n_scenes, steps=1500, every=50, warmup=50, adam_steps=3000, eval_scenes=2,
                              save_path=None):
    """The closure's data set and training at the sizes of its configuration:
    coarse 96x48x48 against 2x-refined 192x96x96 runs of single-ellipsoid
    channels (semi-axes U[4, 10] coarse cells, random orientation, E
    log-uniform in [2e2, 2e4], nu U[0.2, 0.4], inflow U[0.01, 0.05]; each
    scene at the substep count its material needs), `steps` coarse steps per
    scene, features and targets sampled every `every` coarse steps after
    `warmup`; MLP trained with batch 65,536 and Adam lr 1e-3; then plain and
    closure coarse runs of held-out scenes against their refined runs over
    the same horizon. Prints progress on stderr."""
    import numpy as np
    import torch

    from fsi.closure import ClosureScene, accuracy_vs_refined, scene_pair_data, train_closure

    scenes = [ClosureScene.random(s, **CFG4_SCENES) for s in range(1000, 1000 + n_scenes)]
    t0 = time.perf_counter()
    Xs, Ts = [], []
    for i, sc in enumerate(scenes):
        X, T = scene_pair_data(sc, steps, every=every, warmup=warmup)
        Xs.append(X)
        Ts.append(T)
        print(f"[closure data] scene {i + 1}/{n_scenes}: {len(X)} rows, {sc.substeps} substeps, "
              f"{time.perf_counter() - t0:.0f} s", file=sys.stderr, flush=True)
    X = np.concatenate(Xs)
    T = np.concatenate(Ts)
    t1 = time.perf_counter()
    closure, hist = train_closure(X, T, steps=adam_steps, batch=65536, lr=1.0e-3, seed=1)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    if save_path:
        closure.save(save_path)
    held = {}
    for s in range(2000, 2000 + eval_scenes):
        sc = ClosureScene.random(s, **CFG4_SCENES)
        held[f"seed {s}"] = {"semi_axes": [round(a, 2) for a in sc.semi_axes], "youngs_modulus": round(sc.youngs_modulus, 1),
                             "inflow": round(sc.inflow, 4), "substeps": sc.substeps,
                             **accuracy_vs_refined(sc, steps, closure, force_every=every, progress=f"held-out {s}")}
    return {"workload": f"closure data: {n_scenes} seeded single-ellipsoid scenes, 96x48x48 coarse vs 192x96x96 "
                        f"refined, {steps} coarse steps each (samples every {every} after {warmup})",
            "substeps": [sc.substeps for sc in scenes],
            "train": {"samples": int(len(X)), "batch": 65536, "lr": 1.0e-3, "adam_steps": adam_steps,
                      "data_s": round(t1 - t0, 1), "train_s": round(t2 - t1, 2),
                      "loss_first": round(hist[0], 4), "loss_last": round(float(np.mean(hist[-50:])), 4)},
            "held_out": held, "closure_file": str(save_path) if save_path else None}


def closure_large_scene(closure_path, steps=2000, speed_steps=2, plain_storages=("i16", "fp32")):
    """The large coupled scene of the closure's configuration: 64 elastic
    spheres (radius 5-14 refined cells, E log-uniform in [5e2, 5e3] per
    sphere, nu 0.3) in a 384x192x192 channel (FP32, plain penalty coupling),
    `steps` coarse steps = 2 x `steps` refined steps, against the 192x96x96
    run with FP16 populations, plain and with the closure saved at
    `closure_path`, and plain with I16 and FP32 populations (storage or
    resolution as the source of the coarse run's error): drag, centroid
    displacement, per-particle force errors, fallback counts, and the time
    per coarse step of the closure run against the refined run. Prints
    progress on stderr."""
    import torch

    from fsi.closure import LearnedClosure, PackScene, accuracy_vs_refined, stable_substeps

    scene = PackScene(shape=(192, 96, 96), count=64, radii=(2.5, 7.0), poisson_ratio=0.3, inflow=0.04,
                      margin=4.0, gap=2.0, seed=5, body_modulus=(5.0e2, 5.0e3),
                      substeps=stable_substeps(5.0e3, 0.3, 1.2))
    closure = LearnedClosure.load(closure_path) if closure_path else None
    acc = accuracy_vs_refined(scene, steps, closure, storage="fp16", force_every=50, progress="large scene",
                              plain_storages=plain_storages)
    coarse = scene.simulation(False, "fp16")
    if closure is not None:
        coarse.coupler.set_closure(closure)
    fine = scene.simulation(True)

    def timed(sim, n):
        sim.step()
        torch.cuda.synchronize()
        a = time.perf_counter()
        for _ in range(n):
            sim.step()
        torch.cuda.synchronize()
        return (time.perf_counter() - a) / n

    tc = timed(coarse, speed_steps)
    tf = timed(fine, 2 * speed_steps) * 2.0
    return {"workload": f"64 spheres (r 5-14 refined cells, E log-uniform [5e2, 5e3] per sphere), 384x192x192 "
                        f"FP32 refined ({2 * steps} steps) vs 192x96x96 FP16 coarse ({steps} steps), "
                        f"{scene.substeps} substeps per lattice step",
            "coarse_particles": int(coarse.mpm.num_particles), "refined_particles": int(fine.mpm.num_particles),
            "accuracy_vs_refined": acc,
            "speed": {"coarse_ms_per_step": round(1e3 * tc, 3), "refined_ms_per_coarse_step": round(1e3 * tf, 3),
                      "speedup": round(tf / tc, 2)}}


def time_coupled(steps, warmup, refine=2):
    import torch

    import fsi

    cfg, solid, pos = coupled_scene_32(refine=refine)
    sim = fsi.FSISimulation(cfg)
    sim.initialize_lbm(solid)
    sim.initialize_mpm_from_numpy(pos, particle_volume=particle_volume(cfg))
    for _ in range(warmup):
        sim.advance()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        sim.advance()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return {"steps_per_s": round(steps / dt, 3), "ms_per_step": round(1e3 * dt / steps, 4),
            "particles": int(len(pos)), "substeps": cfg.coupling.mpm_substeps_per_lbm_step, "timed_steps": steps}


def calibrate_copy(n):
    """A plain device-to-device copy of one FP32 population buffer
    (19 * n^3 * 4 bytes), three times: the known-size kernel that scales the
    FETCH_SIZE / WRITE_SIZE counters of a profiled run."""
    import torch

    nbytes = 19 * n ** 3 * 4
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "device copy", "bytes_read": nbytes, "bytes_written": nbytes,
                      "launches": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--launch-samples", type=int, default=200)
    ap.add_argument("--coupled-steps", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-coupled", action="store_true")
    ap.add_argument("--no-fp16", action="store_true")
    ap.add_argument("--no-closure", action="store_true")
    ap.add_argument("--no-error-check", action="store_true",
                    help="skip the accuracy runs: the FP64 lattice the FP32/FP16/I16 lattice fields are compared "
                         "with, and the FP32/mixed-precision suspension comparison")
    ap.add_argument("--error-steps", type=int, default=1000,
                    help="steps of the (untimed) lattice accuracy runs against the FP64 build")
    ap.add_argument("--cpu-task", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--cpu-budget", type=float, default=10.0, help=argparse.SUPPRESS)
    ap.add_argument("--large-steps", type=int, default=10,
                    help="timed steps of the 256x128x128 coupled suspension, FP32 and mixed (0: skip)")
    ap.add_argument("--large-error-steps", type=int, default=100,
                    help="coupled steps of the (untimed) suspension runs compared with its FP64 build")
    ap.add_argument("--large-only", action="store_true",
                    help="only the suspension: timing and the FP64 comparison (its own JSON line)")
    ap.add_argument("--large-refine", type=int, default=LARGE_REFINE, choices=(1, 2),
                    help="MPM cells per lattice cell and axis of the timed suspension (2: dx_mpm = dx_lbm / 2)")
    ap.add_argument("--large-error-refine", type=int, default=1, choices=(1, 2),
                    help="the same for the suspension's FP64 comparison")
    ap.add_argument("--closure-config", choices=("train", "large"), default=None,
                    help="only the learned closure at its configuration's sizes (its own JSON line): 'train' "
                         "generates the coarse/refined pair data and trains; 'large' runs the 64-sphere scene")
    ap.add_argument("--closure-scenes", type=int, default=32, help="scenes of --closure-config train")
    ap.add_argument("--closure-steps", type=int, default=0,
                    help="coarse steps per run of --closure-config (default 1500 train, 2000 large)")
    ap.add_argument("--closure-path", default=None, help="closure file written by train and read by large")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain device copy of one population buffer (counter calibration)")
    args = ap.parse_args()
    if args.cpu_task:  # child process of cpu_baseline(): CPU only
        _cpu_task(args.cpu_task, args.cpu_budget)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); multi-GPU sharding "
              "of the lattice is a design note, not part of this build.", file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy:
        calibrate_copy(args.size)
        return
    if args.closure_config == "train":
        out = closure_data_and_training(args.closure_scenes, steps=args.closure_steps or 1500,
                                        save_path=args.closure_path)
        print(json.dumps(out), flush=True)
        return
    if args.closure_config == "large":
        print(json.dumps(closure_large_scene(args.closure_path, steps=args.closure_steps or 2000)), flush=True)
        return
    if args.large_only:
        r = args.large_refine
        out = {"workload": large_workload(r)}
        if args.large_steps > 0:
            out["fp32"] = time_coupled_large(args.large_steps, refine=r)
            out["mixed_i16"] = time_coupled_large(args.large_steps, "i16", refine=r)
        if args.large_error_steps > 0:
            out["error_vs_fp64"] = {"workload": large_workload(args.large_error_refine),
                                    **large_errors_vs_fp64(args.large_error_steps, refine=args.large_error_refine)}
        print(json.dumps(out), flush=True)
        return

    import numpy as np
    import torch

    import fsi
    from fsi.scenes import porous_spheres

    n = args.size
    mask, info = porous_spheres((n, n, n), porosity=0.8, rmin=4.0, rmax=10.0, seed=0)
    n_cells = mask.size
    n_fluid = int((mask == 0).sum())

    runs = {}
    fields = {}
    line_errors = None
    storages = ["fp32"] if args.no_fp16 else ["fp32", "fp16", "i16"]
    for storage in storages:
        s = fsi.LBMSolver3D(periodic_lbm_config(n, storage))
        s.initialize(mask)
        wall, launch_ms = time_lattice(s, args.steps, args.warmup, args.launch_samples)
        b = 4 if storage == "fp32" else 2
        alg = algorithmic_bytes(n_fluid, n_cells, b)
        runs[storage] = {
            "mlups": n_cells * args.steps / wall / 1e6,
            "mflups": n_fluid * args.steps / wall / 1e6,
            "ms_per_step": 1e3 * wall / args.steps,
            "launch_ms": launch_ms,
            "bytes": alg,
            "achieved_gbs": alg / (launch_ms * 1e-3) / 1e9,
        }
        del s
        torch.cuda.empty_cache()
        if not args.no_error_check:
            # untimed accuracy run of exactly --error-steps steps from initialize()
            s = fsi.LBMSolver3D(periodic_lbm_config(n, storage))
            s.initialize(mask)
            for i in range(args.error_steps):
                s.step()
            s._ensure_macro()
            extra = {"storage_range_used": round(s.storage_range(), 4)} if storage == "i16" else {}
            fields[storage] = (s.rho.clone(), s.vel.clone(), extra)
            del s
            torch.cuda.empty_cache()
    if fields:
        line_errors = lattice_errors_vs_fp64(n, mask, fields, args.error_steps)

    r32 = runs["fp32"]
    traffic = read_traffic(f"lbm_step_{n}_fp32_pmc.txt")
    line = {
        "metric": METRIC,
        "value": round(r32["mlups"], 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r32["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"D3Q19 MRT lattice {n}^3, porous sphere mask (porosity "
                        f"{info['porosity']:.4f}, r 4-10, seed 0), periodic, g_x=1e-6, tau=0.55, "
                        "FP32 populations",
            "cells": n_cells,
            "fluid_cells": n_fluid,
            "kernel": "k_lbm_pull<float> stream+collide (fused pull-scheme MRT step)",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(r32["achieved_gbs"], 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(r32["achieved_gbs"] / HBM_PEAK_GBS, 4),
            "traffic": traffic,
        },
        "mflups": round(r32["mflups"], 2),
        "launch_ms": round(r32["launch_ms"], 5),
    }
    for storage, label in (("fp16", "IEEE binary16 of (f - w rho0) 2^15"),
                           ("i16", "16-bit integers of (f - w rho0)/w with an adaptive scale")):
        if storage not in runs:
            continue
        r16 = runs[storage]
        line[storage] = {
            "storage": label,
            "mlups": round(r16["mlups"], 2),
            "ms_per_step": round(r16["ms_per_step"], 5),
            "launch_ms": round(r16["launch_ms"], 5),
            "roofline_frac": round(r16["achieved_gbs"] / HBM_PEAK_GBS, 4),
            "achieved_gbs": round(r16["achieved_gbs"], 1),
            "speedup_vs_fp32": round(r16["mlups"] / r32["mlups"], 3),
            "traffic": read_traffic(f"lbm_step_{n}_{storage}_pmc.txt"),
        }
    if line_errors is not None:
        line["error_vs_fp64"] = line_errors
    if not args.no_coupled:
        line["coupled"] = {"workload": "32^3 coupled penalty scene, FP32, 4096 particles, MPM grid 64^3 at "
                                       "dx_mpm = dx_lbm / 2, 224 substeps",
                           **time_coupled(args.coupled_steps, 2)}
        line["coupled_colocated"] = {"workload": "the same cube on a co-located 32^3 MPM grid (twice the "
                                                 "particle spacing), 112 substeps",
                                     **time_coupled(args.coupled_steps, 2, refine=1)}
    if args.large_steps > 0:
        r = args.large_refine
        line["coupled_large"] = {"workload": large_workload(r) + ", FP32",
                                 **time_coupled_large(args.large_steps, refine=r)}
        # mixed precision: I16 lattice populations, FP32 MPM and coupling
        line["coupled_large_mixed"] = {
            "workload": "the same suspension, I16 lattice storage, FP32 MPM and coupling",
            **time_coupled_large(args.large_steps, storage="i16", refine=r)}
        line["coupled_large_mixed"]["speedup_vs_fp32"] = round(
            line["coupled_large_mixed"]["steps_per_s"] / line["coupled_large"]["steps_per_s"], 3)
        if r != 1:
            line["coupled_large_colocated"] = {"workload": large_workload(1) + ", FP32",
                                               **time_coupled_large(args.large_steps, refine=1)}
        if not args.no_error_check and args.large_error_steps > 0:
            line["coupled_large_error_vs_fp64"] = {
                "workload": large_workload(args.large_error_refine),
                **large_errors_vs_fp64(args.large_error_steps, refine=args.large_error_refine)}
    if not args.no_closure:
        line["learned_closure"] = time_closure()
    if not args.no_cpu_baseline:
        line["cpu_baseline"], line["cpu_baseline_more"] = cpu_baseline()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
