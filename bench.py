"""Timestep benchmark of the HIP hot path.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256]

A step is one frame of SIMULATION_WORLD::AdvanceOneFrame on the oil/water
rising-drop scene at n^3 (default 256^3): the CFL rule's dt for the current
velocity (recomputed before every frame, SIMULATION_WORLD.h:288-312),
level-set WENO5 stages, ENO3 MAC advection, gravity, explicit viscosity, the
variable-density projection solved to 1e-8 relative residual, and 2 x 25
reinitialization sweeps.  The default run is the workload's 50 frames from
rest (5 untimed, 45 timed).  The headline value is ms/step with the FP32
inner solve inside FP64 iterative refinement, the inner solve being CG
preconditioned by a geometric multigrid V-cycle, and the WENO5 derivatives
of the advection and reinitialization in the fewer-operation form
(fast_weno, FP64; within rounding of the reference's order); the same path
in the reference's operation order, FP32 Jacobi-PCG inside the same
refinement and the plain FP64 Jacobi-PCG path run the same frames from the
same state for comparison (`--paths`), each on its own, so that one
path's failure (named by frame, dt and residual) leaves the others'
figures.  Inputs are resident on the GPU before the timed region; every step
is synchronous.

One JSON line is printed.  `roofline` is for the dominant kernel (timed with
HIP events on the solver's stream); `cpu_baseline` times the CPU checker on
the host: two complete frames of the same workload with the PCG capped
(see `sample`), and the 64^3 dam break's 20 frames end to end on the CPU
and on the GPU (`c0_end_to_end`).

Other workloads (each its own JSON line): `--workload batch` (48 uneven
scenes advanced together against one at a time), `--workload adjoint` (the
checkpointed rollout gradient against central differences), `--workload
projection` (one projection of the 64-bubble pool, density ratio 1000, at
512^3 per path), `--workload pipe` (the reference scripts' oil/water pipe
scenes: 3D_script.sim's parallel pipe at 51^3 by default, `--pipe-script
3D_script_1.sim` or `3D_script_3.sim` the vertical pipe at 71 x 51 x 71).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=45)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=None, help="grid edge (default 256; 128 for the adjoint workload)")
    ap.add_argument("--scene", default="rising_drop")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-c0", action="store_true",
                    help="skip the end-to-end CPU/GPU comparison on the 64^3 dam break (20 frames)")
    ap.add_argument("--paths", default="mg-fast,mg,refine,fp64",
                    help="projection paths run on the same frames; the first is the headline: mg (FP32 CG with a "
                         "multigrid preconditioner inside FP64 refinement), refine (FP32 Jacobi-PCG inside FP64 "
                         "refinement), fp64 (FP64 Jacobi-PCG); a -fast suffix takes the WENO5 derivatives in "
                         "fewer operations (fast_weno)")
    ap.add_argument("--no-fp64", action="store_true", help="drop fp64 from --paths")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a 1 GiB device copy first (profiler byte-counter calibration)")
    ap.add_argument("--workload", default="timestep",
                    choices=["timestep", "batch", "adjoint", "projection", "sensitivity", "pipe"],
                    help="batch: 48 uneven scenes advanced together vs one at a time; adjoint: checkpointed "
                         "rollout gradient checked against central differences; projection: one projection of "
                         "the 64-bubble pool (default 512^3) per path; pipe: the oil/water pipe scene (default "
                         "51^3) (each its own JSON line)")
    ap.add_argument("--max-iteration", type=int, default=40000, help="projection workload: PCG iteration cap")
    ap.add_argument("--batch-path", default="mg", choices=["mg", "refine"],
                    help="batch workload: the FP32 projection inside FP64 refinement (multigrid or Jacobi)")
    ap.add_argument("--directions", type=int, default=32)
    ap.add_argument("--rollout-steps", type=int, default=16)
    ap.add_argument("--checkpoint-every", type=int, default=4)
    ap.add_argument("--density-ratio", type=float, default=1000.0,
                    help="adjoint workload: water/air density ratio (water 1000 kg/m^3)")
    ap.add_argument("--scenes", type=int, default=48)
    ap.add_argument("--pipe-script", default="3D_script.sim", choices=["3D_script.sim", "3D_script_1.sim", "3D_script_3.sim"],
                    help="pipe workload: the reference script whose pipe runs (3D_script.sim the parallel pipe at "
                         "n^3, default 51; 3D_script_1.sim and 3D_script_3.sim the vertical pipe at 71 x 51 x 71)")
    ap.add_argument("--threads", type=int, default=0)
    return ap.parse_args()


def measured_traffic(scene, n, kernel):
    """HBM bytes per launch of `kernel` (FETCH_SIZE and WRITE_SIZE per
    dispatch from rocprofv3 PMC runs of this bench, scaled by the counters'
    reading of a 1 GiB copy), when a profile of this scene and size is
    committed under profiles/; else None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{scene}_{n}.json")
    try:
        with open(path) as f:
            k = json.load(f)["kernels"][kernel]
        return int(k["read_bytes"] + k["write_bytes"])
    except (OSError, KeyError, ValueError):
        return None


# VALU issue peak of the MI355X: 256 CUs x 4 SIMDs, one 64-lane FP32 or
# FP64 instruction per SIMD every 4 cycles at 2.4 GHz
VALU_PEAK_WAVE_INSTR_PER_S = 256 * 4 * 2.4e9 / 4


def measured_valu(scene, n, kernel, avg_ms):
    """VALU issue rate of `kernel` from its SQ_INSTS_VALU / SQ_WAVES counts
    per dispatch (rocprofv3 PMC run of this bench, committed under
    profiles/) over the live average launch time; None without a profile."""
    path = os.path.join(ROOT, "profiles", f"traffic_{scene}_{n}.json")
    try:
        with open(path) as f:
            k = json.load(f)["kernels"][kernel]
        ipw, waves = float(k["valu_instructions_per_wave"]), float(k["waves_per_dispatch"])
    except (OSError, KeyError, ValueError):
        return None
    rate = ipw * waves / (avg_ms * 1e-3)
    out = {"instructions_per_wave": ipw, "waves": waves, "achieved_wave_instr_per_s": round(rate, -6),
           "peak_wave_instr_per_s": VALU_PEAK_WAVE_INSTR_PER_S, "frac": round(rate / VALU_PEAK_WAVE_INSTR_PER_S, 4)}
    if kernel.startswith("k_reinit"):
        f = reinit_cu_fraction(kernel == "k_reinit_zm")
        out["reinit_cu_fraction"] = f
        out["frac_of_its_cus"] = round(rate / (VALU_PEAK_WAVE_INSTR_PER_S * f), 4)
    return out


def reinit_cu_fraction(fast=False):
    """Fraction of the CUs the step's reinitialization stream may use (the
    solver's default 3/4, 5/8 with fast WENO5, FS_REINIT_CU_FRAC; all of them
    when the step runs in sequence, FS_SERIAL_REINIT)."""
    if os.environ.get("FS_SERIAL_REINIT"):
        return 1.0
    try:
        f = float(os.environ.get("FS_REINIT_CU_FRAC", "0.625" if fast else "0.75"))
    except ValueError:
        return 1.0
    return f if 0 < f < 1 else 1.0


def divergent_velocity(n):
    """Smooth MAC velocity with nonzero divergence, so that a projection on
    its own has work to do (the scenes start at rest)."""
    c = (np.arange(n) + 0.5) / n
    f = np.arange(n + 1) / n
    u = 0.1 * np.sin(np.pi * f)[None, None, :] * np.cos(np.pi * c)[None, :, None] * np.ones((n, 1, 1))
    v = 0.1 * np.sin(2 * np.pi * f)[None, :, None] * np.cos(np.pi * c)[:, None, None] * np.ones((1, 1, n))
    w = 0.1 * np.sin(3 * np.pi * f)[:, None, None] * np.cos(np.pi * c)[None, None, :] * np.ones((1, n, 1))
    return u, v, w


def cpu_frames_timed(lib, scene, n, frames, max_iteration=20000, allow_cap=False):
    """Frames of the CPU checker from the scene's initial state through the
    shared frame loop (dt recomputed per frame, CFL 0.5, |g|), each timed end
    to end; with allow_cap, a projection stopped by max_iteration is kept.
    Returns (seconds per frame, records)."""
    from fluid_simulator_backup_amd import _abi, frame_loop, scenes
    phi, vel, P, g = scenes.make(scene, n, stop_rule=_abi.FS_STOP_REL_RESIDUAL, rel_tolerance=1e-8,
                                 max_iteration=max_iteration)
    s = _abi.Solver(lib, "fso_", P, g)
    lib.fso_set_viscosity_inplace(s.h, 0)  # the HIP path's out-of-place order
    s.set_fields(phi, *vel)
    secs, recs = [], []
    for f in range(frames):
        dt = s.cfl_dt(0.5, 1)
        t = time.perf_counter()
        rc, st = frame_loop.advance_frame_rc(s, dt)
        secs.append(time.perf_counter() - t)
        if rc != _abi.FS_OK and not (allow_cap and rc == _abi.FS_ERR_NOT_CONVERGED):
            s.close()
            raise frame_loop.FrameError(f, dt, rc, "CPU checker", st)
        recs.append(dict(st, dt=dt))
    s.close()
    return secs, recs


def cpu_baseline(scene, n, iters_per_step, threads, cap=(100, 300)):
    """Per-step time of the CPU checker (C, OpenMP, FP64, CSR Jacobi-PCG as
    the reference) on the headline workload, from two complete frames timed
    end to end: the scene's first frame with the PCG capped at cap[0] and
    again at cap[1] iterations.  Their difference is the per-iteration cost;
    the per-step value adds the FP64 path's remaining iterations at it (a
    full 256^3 CPU frame to 1e-8 runs for minutes)."""
    from fluid_simulator_backup_amd import _abi
    lib = _abi.load(os.path.join(ROOT, "oracle", "libfs_oracle.so"), "fso_")
    lib.fso_set_threads(threads)
    t = {}
    for c in cap:
        secs, _ = cpu_frames_timed(lib, scene, n, 1, max_iteration=c, allow_cap=True)
        t[c] = secs[0]
        print(f"[bench] cpu baseline: one {n}^3 frame, PCG capped at {c}: {t[c]:.2f} s", file=sys.stderr, flush=True)
    t_iter = (t[cap[1]] - t[cap[0]]) / (cap[1] - cap[0])
    per_step = t[cap[0]] + max(iters_per_step - cap[0], 0) * t_iter
    return {
        "value": round(per_step * 1e3, 1), "unit": "ms/step", "cores": threads, "kind": "port",
        "sample": (f"CPU checker (C, OpenMP, FP64, CSR Jacobi-PCG) with {threads} threads: the {scene} {n}^3 "
                   f"first frame timed end to end twice, PCG capped at {cap[0]} ({t[cap[0]]:.2f} s) and at "
                   f"{cap[1]} ({t[cap[1]]:.2f} s) iterations -> {t_iter * 1e3:.1f} ms/iteration; value = the "
                   f"{cap[0]}-iteration frame + the FP64 GPU path's remaining {max(iters_per_step - cap[0], 0):.0f} "
                   f"iterations per frame at that cost (a model of the full frame from measured frames); "
                   f"{t[cap[0]] + t[cap[1]]:.1f} s timed"),
        "frame_capped_s": {str(c): round(t[c], 3) for c in cap},
        "ms_per_pcg_iteration": round(t_iter * 1e3, 3),
    }


def c0_comparison(lib_gpu, threads, frames=20, n=64):
    """The reference's CPU-runnable case measured end to end on both sides:
    the air/water dam break at 64^3, `frames` frames of the plain FP64
    method (WENO5 level set with reinitialization, ENO3 MAC advection,
    explicit viscosity, Jacobi-PCG to 1e-8 relative residual) on the CPU
    checker, and the same frames on the GPU with FP64 and with the headline
    multigrid path."""
    from fluid_simulator_backup_amd import _abi, frame_loop, scenes
    lib = _abi.load(os.path.join(ROOT, "oracle", "libfs_oracle.so"), "fso_")
    lib.fso_set_threads(threads)
    secs, recs = cpu_frames_timed(lib, "dam_break", n, frames)
    out = {"workload": f"dam_break {n}^3 (air/water 1/1000), {frames} frames from rest, dt from the CFL rule each "
                       f"frame (0.5, |g|), projection to 1e-8 relative residual",
           "cpu": {"ms_per_step": round(1e3 * sum(secs) / frames, 2), "cores": threads,
                   "pcg_iterations_per_step": float(np.mean([r["pcg_iterations"] for r in recs]))}}
    print(f"[bench] c0 dam break {n}^3 CPU checker: {out['cpu']['ms_per_step']:.1f} ms/step", file=sys.stderr,
          flush=True)
    import torch
    for path in ("fp64", "mg"):
        phi, vel, P, g = scenes.make("dam_break", n, stop_rule=_abi.FS_STOP_REL_RESIDUAL, rel_tolerance=1e-8,
                                     max_iteration=20000, precision=getattr(_abi, PATHS[path]))
        s = _abi.Solver(lib_gpu, "fs_", P, g, 0)
        s.set_fields(phi, *vel)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        grecs = frame_loop.advance_frames(s, frames)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / frames
        f = s.get_fields()
        s.close()
        out[path] = {"ms_per_step": round(ms, 3), "speedup_vs_cpu": round(out["cpu"]["ms_per_step"] / ms, 2),
                     "pcg_iterations_per_step": float(np.mean([r["pcg_iterations"] for r in grecs])),
                     "max_residual_rel": max(r["residual_rel"] for r in grecs)}
        print(f"[bench] c0 dam break {n}^3 GPU {LABEL[path]}: {ms:.2f} ms/step", file=sys.stderr, flush=True)
        del f
    return out


def batch_main(a):
    """Uneven scenes (edges 32..128, dam break or rising drop, density ratio
    1.25..1000) advanced together through fs_batch_advance_frame, against
    the same scenes advanced one at a time; the FP32 projection inside FP64
    refinement (--batch-path: multigrid or Jacobi) to 1e-8."""
    import torch

    import fluid_simulator_backup_amd as fsa
    from fluid_simulator_backup_amd import _abi, frame_loop, scenes
    torch.cuda.set_device(0)
    lib = fsa.load_library()
    prec = getattr(_abi, PATHS[a.batch_path])
    sc = scenes.batch_scenes(a.scenes, 3, precision=prec, stop_rule=_abi.FS_STOP_REL_RESIDUAL,
                             rel_tolerance=1e-8, max_iteration=20000)
    cells = sum(g.nx * g.ny * g.nz for _, _, _, _, g in sc)
    batch = _abi.Batch(lib, [s[3] for s in sc], [s[4] for s in sc], 0, a.threads)
    for (_, phi, vel, _, _), bs in zip(sc, batch.scenes):
        bs.set_fields(phi, *vel)
    print(f"[bench] batch of {len(sc)} scenes, {cells / 1e6:.1f} M cells, warmup={a.warmup} steps={a.steps}",
          file=sys.stderr, flush=True)

    def frame_batch():
        # every scene's dt from the CFL rule on its own velocity, each frame
        dts = [bs.cfl_dt(0.5, 1) for bs in batch.scenes]
        return sum(st["pcg_iterations"] for st in batch.advance_frame(dts))

    for _ in range(a.warmup):
        frame_batch()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    its = 0
    for _ in range(a.steps):
        its += frame_batch()
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    batch.close()
    print(f"[bench] batch {t_batch * 1e3 / a.steps:.1f} ms/frame", file=sys.stderr, flush=True)
    # the same scenes, one at a time
    t_seq = 0.0
    its_seq = 0
    for (_, phi, vel, P, g) in sc:
        s = _abi.Solver(lib, "fs_", P, g, 0)
        s.set_fields(phi, *vel)
        frame_loop.advance_frames(s, a.warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        its_seq += sum(r["pcg_iterations"] for r in frame_loop.advance_frames(s, a.steps))
        torch.cuda.synchronize()
        t_seq += time.perf_counter() - t0
        s.close()
    print(f"[bench] one at a time {t_seq * 1e3 / a.steps:.1f} ms/frame", file=sys.stderr, flush=True)
    out = {
        "metric": "aggregate cells/s, uneven scenes advanced together", "value": round(cells * a.steps / t_batch, 1),
        "unit": "cells/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(t_batch * 1e3 / a.steps, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32-inner/f64-refined (pressure solve), f64 elsewhere", "data": "synthetic",
        "config": {"workload": f"{len(sc)} scenes, edges 32-128, dam break / rising drop, density ratio 1.25-1000, "
                               f"{LABEL[a.batch_path]} to 1e-8, host threads {a.threads or min(len(sc), 16)}",
                   "cells": cells},
        "one_at_a_time_cells_per_s": round(cells * a.steps / t_seq, 1),
        "speedup_vs_one_at_a_time": round(t_seq / t_batch, 3),
        "pcg_iterations_per_step": its / a.steps, "pcg_iterations_per_step_one_at_a_time": its_seq / a.steps,
    }
    print(json.dumps(out))


def adjoint_main(a):
    """Adjoint rollout on the sloshing tank (water below z = 0.5 + 0.1 (x -
    0.5), air 1/1.8e-5, water 1000/1e-3) at n^3: `rollout_steps` FP64 steps
    of the fixed schedule (WENO5 level set and velocity, explicit viscosity,
    projection to 1e-12 relative residual, 2 x 25 reinitialization sweeps),
    loss sum (H(phi_T) - H(phi_target))^2 with eps = 1.5 dx.  phi_target is
    the end of a run from the Gaussian random velocity of seed 4 (sigma 8
    cells, std 0.05 m/s); the gradient is taken at the one of seed 5 with a
    checkpoint every `checkpoint_every` steps, and checked along the
    unit-RMS Gaussian random directions of seeds 40 + d, d < directions,
    against central differences with step 1e-6 and their Richardson
    extrapolation with step 5e-7 (which removes the h^2 term)."""
    import torch

    import fluid_simulator_backup_amd as fsa
    from fluid_simulator_backup_amd import _abi, scenes
    torch.cuda.set_device(0)
    lib = fsa.load_library()
    n = a.n
    phi, _, P, g = scenes.make("slosh", n, schedule=_abi.FS_SCHEDULE_FIXED, use_5th_weno_v=1,
                               stop_rule=_abi.FS_STOP_REL_RESIDUAL, rel_tolerance=1e-12, max_iteration=50000,
                               air_density=1000.0 / a.density_ratio)
    s = _abi.Solver(lib, "fs_", P, g, 0)
    zp = np.zeros_like(phi)

    def set_state(vel):
        s.set_fields(phi, *vel, p=zp)

    def forward(vel):
        set_state(vel)
        for _ in range(a.rollout_steps):
            # 1e-12 is below FP64's attainable residual at ratio 1000: the
            # solve stops there and reports it, the step goes on
            s.step(dt, allow_not_converged=True)

    sig = 8.0 * n / 128
    v_target = scenes.gaussian_random_velocity(n, 4, sigma_cells=sig)
    v0 = scenes.gaussian_random_velocity(n, 5, sigma_cells=sig)
    set_state(v0)
    dt = s.cfl_dt(0.5, 1)
    forward(v_target)
    target = s.get_fields()["phi"]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    forward(v0)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    loss0 = s.levelset_loss(target)
    set_state(v0)
    t0 = time.perf_counter()
    loss, grads, st = s.adjoint_rollout(a.rollout_steps, dt, a.checkpoint_every, target,
                                        allow_not_converged=True)
    torch.cuda.synchronize()
    t_adj = time.perf_counter() - t0
    print(f"[bench] adjoint {n}^3 steps={a.rollout_steps} dt={dt:.4e} loss={loss:.6e} forward {t_fwd:.2f} s, "
          f"gradient {t_adj:.2f} s", file=sys.stderr, flush=True)
    h = 1e-6
    rels, rels_plain = [], []

    def central(d, t):
        forward([x + t * y for x, y in zip(v0, d)])
        lp = s.levelset_loss(target)
        forward([x - t * y for x, y in zip(v0, d)])
        return (lp - s.levelset_loss(target)) / (2 * t)

    for q in range(a.directions):
        d = scenes.gaussian_random_velocity(n, 40 + q, sigma_cells=sig, std=1.0)
        f1, f2 = central(d, h), central(d, h / 2)
        rich = (4 * f2 - f1) / 3
        ad = sum(float(np.vdot(gq, dq)) for gq, dq in zip(grads[1:], d))
        rels.append(abs(rich - ad) / abs(ad))
        rels_plain.append(abs(f1 - ad) / abs(ad))
        print(f"[bench] direction {q}: central(1e-6) {f1:.10e} richardson {rich:.10e} adjoint {ad:.10e} "
              f"rel {rels[-1]:.2e} (central alone {rels_plain[-1]:.2e})", file=sys.stderr, flush=True)
    out = {
        "metric": "adjoint rollout gradient vs central differences (Richardson, steps 1e-6 and 5e-7)",
        "value": float(max(rels)),
        "unit": "max relative error", "n_gpus": 1, "steps": a.rollout_steps, "warmup": 0,
        "ms_per_step": round(t_adj * 1e3 / a.rollout_steps, 3), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"slosh {n}^3 (density ratio {a.density_ratio:g}), {a.rollout_steps} FP64 "
                               f"fixed-schedule steps, checkpoint every "
                               f"{a.checkpoint_every}, {a.directions} unit-RMS Gaussian random directions",
                   "cells": n ** 3, "dt": dt},
        "loss": loss, "loss_forward_only": loss0, "median_rel_err": float(np.median(rels)),
        "central_1e-6_alone_max_rel_err": float(max(rels_plain)),
        "central_1e-6_alone_median_rel_err": float(np.median(rels_plain)),
        "forward_rollout_s": round(t_fwd, 3), "gradient_rollout_s": round(t_adj, 3),
        "gradient_over_forward": round(t_adj / t_fwd, 3), "pcg_iterations_gradient_rollout": st["pcg_iterations"],
        "target": "every direction within 1e-5",
        "directions_within_1e-5": int(sum(r <= 1e-5 for r in rels)),
    }
    print(json.dumps(out))


def sensitivity_main(a):
    """The headline scene's sensitivity to rounding: the same 50 rising-drop
    frames run twice on the plain FP64 path, with two tilings of the
    Jacobi-PCG kernels (FS_PCG_TILING), which differ only in the order of the
    FP64 partial sums; per-field differences after selected frames, in the
    measure the FP32 paths' error against FP64 is reported in."""
    import torch

    import fluid_simulator_backup_amd as fsa
    from fluid_simulator_backup_amd import _abi, frame_loop, scenes
    torch.cuda.set_device(0)
    lib = fsa.load_library()
    n = a.n or 256
    frames = a.warmup + a.steps
    snap_at = sorted({f for f in (1, 5, 10, 20, 30, 40, 50, frames) if f <= frames})
    tilings = ("8,256,8,256", "64,2048,16,1024")
    runs = []
    for tiling in tilings:
        os.environ["FS_PCG_TILING"] = tiling
        phi, vel, P, g = scenes.make(a.scene, n, stop_rule=_abi.FS_STOP_REL_RESIDUAL, rel_tolerance=1e-8,
                                     precision=_abi.FS_FP64, max_iteration=20000)
        s = _abi.Solver(lib, "fs_", P, g, 0)
        s.set_fields(phi, *vel)
        snaps = {}

        def keep(r):
            if r["frame"] + 1 in snap_at:
                snaps[r["frame"] + 1] = s.get_fields()
                print(f"[bench] tiling {tiling}: frame {r['frame'] + 1}", file=sys.stderr, flush=True)

        frame_loop.advance_frames(s, frames, on_frame=keep)
        s.close()
        runs.append(snaps)
    os.environ.pop("FS_PCG_TILING", None)
    diff = {f"after_frame_{f}": field_error(runs[1][f], runs[0][f]) for f in snap_at}
    print(json.dumps({
        "metric": "FP64 run-to-run difference from the order of the PCG partial sums alone", "value":
            diff[f"after_frame_{snap_at[-1]}"]["phi"]["rel_l2"], "unit": "relative L2 of phi after the last frame",
        "n_gpus": 1, "steps": frames, "warmup": 0, "higher_is_better": False, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{a.scene} {n}^3, {frames} frames from rest, CFL dt per frame, FP64 Jacobi-PCG to "
                               f"1e-8, Jacobi tilings {tilings[0]} and {tilings[1]} (zc,nb of the SpMV, of the "
                               f"update)", "cells": n ** 3},
        "difference": diff}))


def projection_main(a):
    """Ill-conditioned projection: the 64-bubble pool (water z < 0.7 with 64
    seeded air bubbles, radii U(0.02, 0.06), density ratio 1000) at n^3
    (default 512), a smooth divergent velocity, and one projection per path
    to 1e-8 relative residual.  The fixed schedule corrects every face with
    the matrix's own coefficients, so the corrected field is divergence-free
    to the solve tolerance: max |div u| after the projection is reported
    against before, with iterations, refinement sweeps, time and the inner
    kernels' in-loop rates."""
    import torch

    import fluid_simulator_backup_amd as fsa
    from fluid_simulator_backup_amd import _abi, scenes
    torch.cuda.set_device(0)
    lib = fsa.load_library()
    n = a.n or 512
    paths = [p for p in a.paths.split(",") if p]
    vel = divergent_velocity(n)

    def div_max(u, v, w):
        d = (u[:, :, 1:] - u[:, :, :-1]) * n
        d += (v[:, 1:, :] - v[:, :-1, :]) * n
        d += (w[1:] - w[:-1]) * n
        return float(np.abs(d).max())

    div0 = div_max(*vel)
    res = {}
    dt = 1e-3
    t0 = time.perf_counter()
    phi, _, P0, g = scenes.make("bubbles", n, stop_rule=_abi.FS_STOP_REL_RESIDUAL, rel_tolerance=1e-8,
                                max_iteration=a.max_iteration, schedule=_abi.FS_SCHEDULE_FIXED)
    print(f"[bench] bubbles {n}^3 level set built in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    for path in paths:
        prec = getattr(_abi, PATHS[path])
        P = type(P0).from_buffer_copy(P0)
        P.precision = prec
        s = _abi.Solver(lib, "fs_", P, g, 0)
        s.set_fields(phi, *vel)
        _abi.set_kernel_sampling(s, True)
        print(f"[bench] projection bubbles {n}^3 {LABEL[path]}", file=sys.stderr, flush=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = s.projection(dt, allow_not_converged=True)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        r = {"projection_s": round(t, 3), "pcg_iterations": st["pcg_iterations"],
             "refinement_sweeps": st["refinement_sweeps"], "residual_rel": st["residual_rel"],
             "converged": st["residual_rel"] <= 1e-8}
        cells = n ** 3
        if path == "mg":
            avg, ns, launches = _abi.frame_kernel_samples(s, _abi.FS_KS_MG_SMOOTH)
            r["finest_sweep_ms"] = round(avg, 4)
            r["finest_sweep_gbs"] = round(6 * 4 * cells / (avg * 1e-3) / 1e9, 1) if avg else None
        else:
            k, cnt = _abi.kernel_samples(s, prec)
            wb = 8 if path == "fp64" else 4
            r["pcg_kernels"] = pcg_roofline(k[:2], cnt, None, cells, n, wb)
        f = s.get_fields()
        s.close()
        r["max_div_after"] = div_max(f["u"], f["v"], f["w"])
        r["div_after_over_before"] = r["max_div_after"] / div0
        del f
        res[path] = r
        print(f"[bench] {LABEL[path]}: {r['projection_s']} s, {r['pcg_iterations']} iterations, "
              f"{r['refinement_sweeps']} sweeps, residual {r['residual_rel']:.2e}, max div {r['max_div_after']:.2e} "
              f"(before {div0:.2e})", file=sys.stderr, flush=True)
    head = paths[0]
    out = {
        "metric": "projection time, 64-bubble pool (density ratio 1000), to 1e-8 relative residual",
        "value": res[head]["projection_s"] * 1e3, "unit": "ms/projection", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": res[head]["projection_s"] * 1e3, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32-inner/f64-refined" if head != "fp64" else "f64", "data": "synthetic",
        "config": {"workload": f"bubbles {n}^3, 64 seeded air bubbles in a water pool z<0.7, density ratio 1000, "
                               f"fixed schedule, one projection of a smooth divergent velocity, dt={dt}",
                   "cells": n ** 3},
        "max_div_before": div0, "paths": res,
    }
    if "fp64" in res:
        for p in paths:
            if p != "fp64":
                res[p]["speedup_vs_fp64"] = round(res["fp64"]["projection_s"] / res[p]["projection_s"], 3)
    print(json.dumps(out))


PATHS = {"mg": "FS_FP32_MG", "refine": "FS_FP32_REFINE", "fp64": "FS_FP64"}
LABEL = {"mg": "fp32+multigrid+refine", "refine": "fp32+jacobi+refine", "fp64": "fp64"}
# "<path>-fast": the same projection path with the WENO5 derivatives in fewer
# operations (fs_params.fast_weno: product-form weights, fused multiply-adds)
for _p in list(PATHS):
    PATHS[_p + "-fast"] = PATHS[_p]
    LABEL[_p + "-fast"] = LABEL[_p] + ", fast WENO5"


def run_path(lib, scene, n, path, warmup, steps, sync):
    """Frames of one projection path from the scene's initial state, each
    with dt recomputed by the CFL rule (SIMULATION_WORLD.h:288-312): `warmup`
    untimed frames, then `steps` timed ones (the CFL reduction and the host's
    read of dt inside the timed region, as in the reference's frame).
    Returns the timing, every frame's record, the in-loop kernel samples and
    the final fields."""
    from fluid_simulator_backup_amd import _abi, frame_loop, scenes
    prec = getattr(_abi, PATHS[path])
    phi, vel, P, g = scenes.make(scene, n, stop_rule=_abi.FS_STOP_REL_RESIDUAL, rel_tolerance=1e-8,
                                 precision=prec, max_iteration=20000, fast_weno=int(path.endswith("-fast")))
    s = _abi.Solver(lib, "fs_", P, g, 0)
    try:
        s.set_fields(phi, *vel)
        del phi, vel
        print(f"[bench] {scene} {n}^3 {LABEL[path]} warmup={warmup} steps={steps}", file=sys.stderr, flush=True)

        def progress(tag):
            def cb(r):
                if r["frame"] % 10 == 9 or r["frame"] == 0:
                    print(f"[bench] {LABEL[path]} {tag} frame {r['frame']}: dt {r['dt']:.4e}, "
                          f"{r['pcg_iterations']} iterations, residual {r['residual_rel']:.2e}, "
                          f"step {r['step_ms']:.1f} ms", file=sys.stderr, flush=True)
            return cb

        snaps = {}
        cb_w = progress("warmup")

        def warm_cb(r):
            cb_w(r)
            # fields after the first frame and at the end of the warmup, for
            # the paths' agreement over time (untimed)
            if r["frame"] + 1 in (1, warmup):
                snaps[r["frame"] + 1] = s.get_fields()

        warm = frame_loop.advance_frames(s, warmup, on_frame=warm_cb)
        sync()
        _abi.set_kernel_sampling(s, True)
        t0 = time.perf_counter()
        recs = frame_loop.advance_frames(s, steps, on_frame=progress("timed"))
        sync()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        print(f"[bench] {LABEL[path]} {ms:.2f} ms/step", file=sys.stderr, flush=True)
        out = {"ms_per_step": ms, "stats": recs, "warmup_stats": warm,
               "frame": {k: _abi.frame_kernel_samples(s, getattr(_abi, k))
                         for k in ("FS_KS_REINIT", "FS_KS_MG_SMOOTH")}}
        if not path.startswith("mg"):
            out["pcg"] = _abi.kernel_samples(s, prec)
            if path.startswith("refine"):
                out["pcg_repeat"] = _abi.time_pcg_kernels(s, prec, 20)
        snaps[warmup + steps] = s.get_fields()
        out["snapshots"] = snaps
        return out
    finally:
        s.close()


def summary(r):
    st = r["stats"]
    allf = r["warmup_stats"] + st
    dts = [s["dt"] for s in allf]
    return {"ms_per_step": round(r["ms_per_step"], 3),
            "pcg_iterations_per_step": float(np.mean([s["pcg_iterations"] for s in st])),
            "refinement_sweeps_per_step": float(np.mean([s["refinement_sweeps"] for s in st])),
            "projection_ms": round(float(np.mean([s["projection_ms"] for s in st])), 3),
            "residual_rel": st[-1]["residual_rel"],
            "max_residual_rel": max(s["residual_rel"] for s in allf),
            "frames": len(allf), "dt_min": min(dts), "dt_max": max(dts),
            "max_cfl_start": max(s["cfl_start"] for s in allf),
            "per_frame": {"dt": [float(f"{d:.6e}") for d in dts],
                          "pcg_iterations": [s["pcg_iterations"] for s in allf],
                          "refinement_sweeps": [s["refinement_sweeps"] for s in allf],
                          "residual_rel": [float(f"{s['residual_rel']:.3e}") for s in allf],
                          "step_ms": [round(s["step_ms"], 2) for s in allf]}}


def field_error(fa, fb):
    """Per-field relative L2 and Linf of fa against fb (pressure with its
    mean removed: the additive constant is held only by the weak pin), and
    the fraction of cells whose phase (phi <= 0, the cell classification)
    agrees."""
    err = {"phase_agreement": float(np.mean((fa["phi"] <= 0) == (fb["phi"] <= 0)))}
    for c in ("phi", "u", "v", "w", "p"):
        a, b = fa[c], fb[c]
        if c == "p":
            a, b = a - a.mean(), b - b.mean()
        d = a - b
        err[c] = {"rel_l2": float(np.linalg.norm(d) / max(np.linalg.norm(b), 1e-300)),
                  "rel_linf": float(np.abs(d).max() / max(np.abs(b).max(), 1e-300))}
    return err


def pcg_roofline(k, cnt, k_rep, cells, n, wb):
    """Jacobi-PCG iteration kernels.  Algorithmic words per cell of one
    iteration (the workload's fixed accounting): SpMV 5 (read p and the
    three face coefficients, write q), update 7 (read x, p, r, q, D^-1,
    write x, r), direction 4 (read r, D^-1, p, write p): 16.  k_pspmv_dot
    does the SpMV and the direction update (9), k_update the update (7).
    Moved words: s, p_old, the level set (or the three coefficient arrays
    below 128^3 or with FS_BETA_ARRAYS), p, q; and x, p, r, q, D, x, r, s."""
    words = {"pspmv_dot": 9, "update": 7}
    phi_coef = n ** 3 >= 128 ** 3 and not os.environ.get("FS_BETA_ARRAYS")
    moved = {"pspmv_dot": 5 if phi_coef else 7, "update": 8}
    names = ["pspmv_dot", "update"]
    gbs = lambda ms, nm, w=words: round(w[nm] * cells * wb / (ms * 1e-3) / 1e9, 1)
    out = {"kernel_ms": {nm: round(float(k[i]), 5) for i, nm in enumerate(names)},
           "samples": {nm: int(cnt[i]) for i, nm in enumerate(names)},
           "algorithmic_words_per_cell": words, "moved_words_per_cell": moved,
           "algorithmic_gbs": {nm: gbs(k[i], nm) for i, nm in enumerate(names)},
           "moved_gbs": {nm: gbs(k[i], nm, moved) for i, nm in enumerate(names)},
           "iteration_algorithmic_gbs": round(16 * cells * wb / (float(k[0] + k[1]) * 1e-3) / 1e9, 1)}
    if k_rep is not None:
        out["back_to_back_repeat_gbs"] = {nm: gbs(k_rep[i], nm) for i, nm in enumerate(names)}
    return out


# Frame kernels sampled in the loop: algorithmic and moved bytes per cell.
# k_reinit, one Sussman sweep in FP64, in the workload's fixed accounting: 6
# words (the ghost copy of phi written and read, the sign field, phi^n read,
# phi^{n+1} written); this build's sweep moves 3 of them (it reads the ghost
# copy and the sign field and writes the next copy's box; a small kernel
# then fills that copy's ghost layers).  k_mg_rb_sweep, one red-black
# Gauss-Seidel sweep of the finest multigrid level in FP32: read x, b and the
# three face-coefficient fields, write x (6 words; the diagonal is formed
# from the coefficients); the same 6 words moved, plus the one-cell ring of
# the first colour's operands that neighbouring tiles also read.
FRAME_KERNELS = {"k_reinit": ("FS_KS_REINIT", 6 * 8, 3 * 8),
                 "k_mg_rb_sweep": ("FS_KS_MG_SMOOTH", 6 * 4, 6 * 4)}
# what bounds each, for the roofline note
LIMITER = {"k_reinit<false>": "FP64 VALU issue: the WENO5 smoothness weights take 24 IEEE divisions per cell and "
                              "every operation keeps the reference's order, so the kernel runs at its instruction "
                              "rate, far below the HBM bound (see the VALU counter summary under profiles/); in the "
                              "frame it runs on a second stream confined to `reinit_cu_fraction` of the CUs, beside "
                              "the projection",
           "k_reinit_zm": "FP64 VALU issue: WENO5 in product-form weights (fast_weno), its smoothness terms "
                          "shared between the two sides and slid along z, still issues several hundred FP64 "
                          "instructions per cell for 48 bytes, far below the HBM bound (see the VALU counter summary "
                          "under profiles/); in the frame it runs on a second stream confined to "
                          "`reinit_cu_fraction` of the CUs, beside the projection",
           "k_mg_rb_sweep": "HBM and load latency of the staged first colour"}


def frame_roofline(R, scene, n, steps, fast=False):
    """Roofline of the multigrid path's dominant sampled frame kernel (the
    larger in-loop time per frame); `fast`: the reinitialization runs the
    fast_weno sweep marching z, k_reinit_zm."""
    cells = n ** 3
    share = {}
    for kn, (ks, _, _) in FRAME_KERNELS.items():
        avg, n_s, launches = R["frame"][ks]
        share[kn] = avg * launches / steps
    kn = max(share, key=share.get)
    ks, alg, mov = FRAME_KERNELS[kn]
    avg, n_s, launches = R["frame"][ks]
    bytes_launch = alg * cells
    achieved = bytes_launch / (avg * 1e-3) / 1e9

    def rates(k):
        a = max(R["frame"][FRAME_KERNELS[k][0]][0], 1e-9)
        return {"avg_launch_ms": round(R["frame"][FRAME_KERNELS[k][0]][0], 5),
                "launches_per_step": R["frame"][FRAME_KERNELS[k][0]][2] / steps,
                "ms_per_step": round(share[k], 3),
                "algorithmic_bytes_per_cell": FRAME_KERNELS[k][1], "moved_bytes_per_cell": FRAME_KERNELS[k][2],
                "algorithmic_gbs": round(FRAME_KERNELS[k][1] * cells / (a * 1e-3) / 1e9, 1),
                "moved_gbs": round(FRAME_KERNELS[k][2] * cells / (a * 1e-3) / 1e9, 1)}

    name = ("k_reinit_zm" if fast else "k_reinit<false>") if kn == "k_reinit" else kn
    return {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(scene, n, name),
            "limiter": LIMITER[name], "valu": measured_valu(scene, n, name, avg),
            "bytes_per_launch": bytes_launch, "moved_bytes_per_launch": mov * cells,
            "moved_frac": round(mov * cells / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "avg_launch_ms": round(avg, 5),
            "timing": f"HIP events on the solver stream around one launch per call (the call's first sweep) of the timed frames "
                      f"({n_s} samples of {launches} launches)",
            "frame_kernels": {k: rates(k) for k in FRAME_KERNELS}}


def jacobi_roofline(R, head, scene, n):
    cells = n ** 3
    k, cnt = R["pcg"]
    k = k[:2]
    wb = 4 if head == "refine" else 8
    pr = pcg_roofline(k, cnt, R.get("pcg_repeat"), cells, n, wb)
    names = ["pspmv_dot", "update"]
    dom = int(np.argmax(k))
    kn = {"pspmv_dot": "k_pspmv_dot", "update": "k_update"}[names[dom]]
    bytes_launch = pr["algorithmic_words_per_cell"][names[dom]] * cells * wb
    achieved = bytes_launch / (k[dom] * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": f"{kn}<{'float' if wb == 4 else 'double'}>", "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(scene, n, kn) if wb == 4 else None,
            "bytes_per_launch": bytes_launch, "avg_launch_ms": round(float(k[dom]), 5),
            "timing": f"HIP events on the solver stream around one iteration per 16 of the timed frames "
                      f"({int(cnt[dom])} samples)"}
    roof.update(pr)
    return roof


def pipe_main(a):
    """The oil/water pipe scene of one of the reference's scripts (core-annular
    flow): 3D_script.sim's parallel pipe at n^3 (default its 51^3, CFL 0.9),
    or the vertical pipe of 3D_script_1.sim / 3D_script_3.sim at their
    71 x 51 x 71 (CFL 0.8); the CFL step before every frame as
    SIMULATION_WORLD does, RK2, the semi-implicit viscosity and the
    projection solved to the script's tolerance.  CPU baseline: the CPU
    checker's frames of the same scene from the same state."""
    import torch

    import fluid_simulator_backup_amd as fsa
    from fluid_simulator_backup_amd import _abi
    torch.cuda.set_device(0)
    lib = fsa.load_library()
    P, Q = _abi.pipe_params(a.pipe_script)
    vertical = Q.is_parallel == 0
    if vertical:
        half, length = (1.8326, 2.618) if a.pipe_script == "3D_script_1.sim" else (2.2270, 3.1814)
        nxz = a.n or 71
        g = _abi.vertical_pipe_grid(nxz, (nxz * 51) // 71, half, length)
        cfl, shape = 0.8, f"{g.nx}x{g.ny}x{g.nz}"
    else:
        n = a.n or 51
        g = _abi.pipe_grid(n)
        cfl, shape = 0.9, f"{n}^3"
    cells = g.nx * g.ny * g.nz
    s = _abi.Pipe(lib, "fs_", P, Q, g)
    s.initial_fields()
    print(f"[bench] pipe {a.pipe_script} {shape} warmup={a.warmup} steps={a.steps}", file=sys.stderr, flush=True)
    stats = []
    for _ in range(a.warmup):
        s.advance_frame(s.cfl_dt(cfl))
    # one launch of each PCG loop kernel per launch batch timed with HIP
    # events on the solver's stream, inside the timed frames
    s.set_sampling(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for f in range(a.steps):
        stats.append(s.advance_frame(s.cfl_dt(cfl)))
        if (f + 1) % 10 == 0:
            print(f"[bench] pipe frame {f + 1}/{a.steps}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    ms = t * 1e3 / a.steps
    ks = {name: s.kernel_samples(k) for name, k in (("k_pcg_spmv_dir", _abi.FS_PIPE_KS_SPMV),
                                                     ("k_pcg_update2", _abi.FS_PIPE_KS_UPDATE))}
    fields = s.get_fields()
    s.close()
    kernels = {}
    for name, (kms, kbytes, nsamp, launches) in ks.items():
        gbs = kbytes / (kms * 1e6) if kms > 0 else None
        kernels[name] = {"avg_us": round(kms * 1e3, 3), "algorithmic_bytes": round(kbytes), "GB/s": gbs and round(gbs, 1),
                         "samples": nsamp, "launches": launches}
    dom = kernels["k_pcg_spmv_dir"]
    out = {
        "metric": "ms/frame, oil/water pipe (core-annular flow), the reference script's scene",
        "value": round(ms, 3), "unit": "ms/step", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(ms, 3), "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (the script's analytic initial state)",
        "config": {"workload": f"oil/water pipe of {a.pipe_script}, {shape}, "
                               f"{'vertical' if vertical else 'parallel'} pipe a = {Q.a}, dimensionless, "
                               f"RK{P.order_for_time_advancing}, CFL {cfl}, projection and semi-implicit viscosity to "
                               f"the script's tolerance 1e-6",
                   "cells": cells},
        "pcg_iterations_per_step": float(np.mean([st["pcg_iterations"] for st in stats])),
        "viscosity_iterations_per_step": float(np.mean([st["viscosity_iterations"] for st in stats])),
        "finite": bool(all(np.all(np.isfinite(v)) for v in fields.values())),
        # the dominant kernel: 120 algorithmic bytes per row of the solved
        # system (7 ELL values and columns, the count, p, z, new p and q)
        "roofline": {"kernel": "k_pcg_spmv_dir", "bound": "hbm", "achieved": dom["GB/s"], "peak": 8000.0,
                     "unit": "GB/s", "frac": dom["GB/s"] and round(dom["GB/s"] / 8000.0, 4), "traffic": None,
                     "rows_per_launch": round(ks["k_pcg_spmv_dir"][1] / 120),
                     "limiter": "launch and load latency of the dependent PCG iterations (two launches each)"},
        "kernels": kernels,
    }
    if not a.no_cpu_baseline:
        orc = _abi.load(os.path.join(ROOT, "oracle", "libfs_oracle.so"), "fso_")
        c = _abi.Pipe(orc, "fsop_", P, Q, g)
        c.initial_fields()
        for _ in range(a.warmup):
            c.advance_frame(c.cfl_dt(cfl))
        frames = max(1, min(a.steps, int(8 * 51 ** 3 / cells) or 1))
        t0 = time.perf_counter()
        for _ in range(frames):
            c.advance_frame(c.cfl_dt(cfl))
        tc = time.perf_counter() - t0
        c.close()
        out["cpu_baseline"] = {"value": round(tc * 1e3 / frames, 1), "unit": "ms/step", "cores": 1, "kind": "port",
                               "sample": f"CPU checker (C, FP64, one thread) of the same scene: frames "
                                         f"{a.warmup + 1}-{a.warmup + frames} after the same warmup, timed end to end"}
        out["speedup_vs_cpu"] = round(tc * 1e3 / frames / ms, 2)
    print(json.dumps(out))


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1 (no multi-GPU path)"}))
        sys.exit(2)
    if a.workload == "batch":
        batch_main(a)
        return
    if a.workload == "adjoint":
        a.n = a.n or 128
        adjoint_main(a)
        return
    if a.workload == "projection":
        projection_main(a)
        return
    if a.workload == "sensitivity":
        sensitivity_main(a)
        return
    if a.workload == "pipe":
        pipe_main(a)
        return
    a.n = a.n or 256
    import torch

    import fluid_simulator_backup_amd as fsa

    torch.cuda.set_device(0)
    lib = fsa.load_library()
    sync = torch.cuda.synchronize
    if a.copy_calibration:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        sync()
        del src, dst
    paths = [p for p in a.paths.split(",") if p]
    if a.no_fp64:
        paths = [p for p in paths if p != "fp64"]
    if not paths or any(p not in PATHS for p in paths):
        sys.exit(f"--paths: a comma-separated subset of {','.join(PATHS)}")
    head = paths[0]
    cells = a.n ** 3
    # each path on its own: a failure is recorded with its frame and residual
    # and does not stop the others
    runs, errors = {}, {}
    for p in paths:
        try:
            runs[p] = run_path(lib, a.scene, a.n, p, a.warmup, a.steps, sync)
        except Exception as e:
            errors[p] = {"error": str(e)}
            if hasattr(e, "frame"):
                errors[p].update({"frame": e.frame, "dt": e.dt,
                                  "residual_rel": e.stats["residual_rel"] if e.stats else None})
            print(f"[bench] {LABEL[p]} failed: {e}", file=sys.stderr, flush=True)
    R = runs.get(head)
    roof = None
    if R is not None:
        try:
            roof = (frame_roofline(R, a.scene, a.n, a.steps, head.endswith("-fast")) if head.startswith("mg")
                    else jacobi_roofline(R, head.split("-")[0], a.scene, a.n))
        except Exception as e:
            roof = {"error": str(e)}
    res = {}
    for p in paths:
        if p in errors:
            res[p] = errors[p]
            continue
        res[p] = summary(runs[p])
        if "pcg" in runs[p]:
            k, cnt = runs[p]["pcg"]
            res[p]["pcg_kernels"] = pcg_roofline(k[:2], cnt, runs[p].get("pcg_repeat"), cells, a.n,
                                                 8 if p.startswith("fp64") else 4)
    if "fp64" in runs:
        for p in runs:
            if p != "fp64":
                res[p]["error_vs_fp64"] = {
                    f"after_frame_{f}": field_error(runs[p]["snapshots"][f], runs["fp64"]["snapshots"][f])
                    for f in sorted(runs[p]["snapshots"]) if f in runs["fp64"]["snapshots"]}
                res[p]["speedup_vs_fp64"] = round(runs["fp64"]["ms_per_step"] / runs[p]["ms_per_step"], 3)
    for r in runs.values():
        r.pop("snapshots", None)

    cpu = None
    if not a.no_cpu_baseline:
        threads = a.cpu_threads or min(16, os.cpu_count() or 1)
        iters = res["fp64"].get("pcg_iterations_per_step") if "fp64" in res else None
        try:
            if iters is None:
                raise RuntimeError("the CPU baseline scales the FP64 path's iteration count: run it (--paths ...,fp64)")
            cpu = cpu_baseline(a.scene, a.n, iters, threads)
        except Exception as e:  # the baseline must not sink the GPU line
            cpu = {"error": str(e)}
        if not a.no_c0:
            try:
                cpu["c0_end_to_end"] = c0_comparison(lib, threads)
            except Exception as e:
                cpu["c0_end_to_end"] = {"error": str(e)}
    h = res.get(head, {})
    ms = R["ms_per_step"] if R is not None else None
    out = {
        "metric": "ms/step at 256^3 two-phase, % HBM peak; FP32+refine vs FP64 at 1e-8 resid",
        "value": round(ms, 3) if ms is not None else None,
        "unit": "ms/step",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms, 3) if ms is not None else None,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-inner/f64-refined (pressure solve), f64 elsewhere" if not head.startswith("fp64") else "f64",
        "data": "synthetic",
        "config": {"workload": f"{a.scene} {a.n}^3, air_bubble_rising (oil 800/0.05 in water 1000/1e-3), "
                               f"{a.warmup} + {a.steps} frames from rest, dt from the CFL rule before every frame "
                               f"(CFL 0.5, |g|), projection to 1e-8 relative residual, {LABEL[head]} ({PATHS[head]})",
                   "cells": cells},
        "path": LABEL[head],
        "cells_per_s": round(cells / (ms * 1e-3), 1) if ms else None,
        "pcg_iterations_per_step": h.get("pcg_iterations_per_step"),
        "refinement_sweeps_per_step": h.get("refinement_sweeps_per_step"),
        "residual_rel": h.get("residual_rel"),
        "max_residual_rel": h.get("max_residual_rel"),
        "dt_range": [h.get("dt_min"), h.get("dt_max")],
        "projection_ms": h.get("projection_ms"),
        "fp64_ms_per_step": res["fp64"].get("ms_per_step") if "fp64" in res else None,
        "speedup_vs_fp64": h.get("speedup_vs_fp64"),
        "roofline": roof,
        "cpu_baseline": cpu,
        "paths": res,
    }
    print(json.dumps(out))
    if errors:
        sys.exit(1)


if __name__ == "__main__":
    main()
