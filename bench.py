#!/usr/bin/env python3
"""Benchmark: DFSPH dam break, 1,000,000 FP32 fluid particles, one MI355X.

A step is one DFSPH solver step (src/DFSPHSolver.cpp:38-112 of the
reference) of the whole scene, inputs resident in HBM.  The default run is
the whole 2000-step dam break after 10 warm-up steps: the density iteration
count per step changes over the run (about 15 during the first hundred steps
of the collapse, about 3 afterwards), so a short window is not
representative.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Prints one JSON line.  value = fluid particle-steps/s over the K timed steps
(host wall clock around the step loop, bracketed by device synchronizes).
The warm-up steps time every kernel with HIP event pairs on the library's
stream (`kernels_warmup`); the dominant one is then the only kernel an event
pair brackets during the timed steps (an event pair costs the stream a few
microseconds: timing every launch would slow the step ~11%).
`roofline.achieved` is the dominant kernel's algorithmic
bytes per launch (per-particle figure x N plus boundary and cell terms)
divided by its average duration per launch that does work (solver loop
sweeps are also launched ahead for converged scenes, which return at once:
their time is charged to the working launches; `avg_launch_us` is the plain
average over all launches, the figure rocprofv3 reports).  `roofline.traffic` is the HBM
traffic per launch of that kernel from the rocprofv3 FETCH_SIZE/WRITE_SIZE
summary in profiles/ (null when that summary is for another kernel), and
`roofline.traffic_frac` that traffic over the same launch duration against
the peak.
cpu_baseline times the FP64 OpenMP restatement of the reference
(oracle/liboracle_fast.so) on the host cores, stepping from the GPU's state
at the end of the timed window.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PMC_SUMMARY = os.path.join(ROOT, "profiles", "dfsph_1m_pmc.json")

# Algorithmic bytes per fluid particle (FP32) of each kernel, and whether the
# launch also counts one boundary sweep (16 B per boundary particle) and one
# cell-table sweep (4 B per cell).  The DFSPH list build also forms the
# divergence solve's initial error: 20 B (rho, alpha) + 36 B (divergence
# error sweep).
BALG = {
    "df_vel_density": (52, 1, 1),
    "df_vel_divergence": (52, 1, 1),
    "df_err_density": (36, 1, 1),
    "df_err_divergence": (36, 1, 1),
    "df_vel_warm": (44, 1, 1),
    "build": (56, 1, 1),
    "dt": (12, 0, 0),
    "df_positions": (40, 0, 0),
    "rank_reorder": (56, 0, 0),
}
KERNELS = ["keys", "scan", "scatter", "rank_reorder", "finalize", "build", "normals", "forces", "dt", "df_predict",
           "df_vel_warm", "df_vel_density", "df_vel_divergence", "df_err_density", "df_err_divergence",
           "df_positions"]


def balg_step(n, nb, ncells, it_d, it_v):
    """Algorithmic HBM bytes of one DFSPH step: 264 B per particle fixed,
    88 B per particle per density and per divergence iteration, 16 B per
    boundary particle for each of the 4 + 2 it_d + 2 it_v boundary sweeps,
    4 B per cell per neighbor sweep and 12 B per cell per grid build."""
    sweeps = 4 + 2 * it_d + 2 * it_v
    return (264 + 88 * (it_d + it_v)) * n + 16 * nb * sweeps + 4 * ncells * sweeps + 12 * ncells


def traffic_for(kernel):
    """HBM bytes per launch of `kernel` from the committed PMC summary, or None."""
    try:
        with open(PMC_SUMMARY) as f:
            pmc = json.load(f)
    except (OSError, ValueError):
        return None
    k = pmc.get("kernels", {}).get(kernel)
    if not k:
        return None
    return int(k["fetch_bytes"] + k["write_bytes"])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n-side", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-roofline", action="store_true", help="no per-kernel events in the timed steps")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the solver has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)

    from sph_amd import configs, sphg

    spec = configs.dfsph_dam_1m(n_side=args.n_side)
    print(f"[bench] scene {spec.name}: {spec.n} fluid, {spec.nb} boundary", file=sys.stderr, flush=True)
    ctx = sphg.Context(0)
    scene = sphg.Scene(ctx, spec.params, spec.fluid, spec.boundary)
    info = scene.info()
    # the warm-up steps time every kernel; the dominant one (largest total
    # time among the kernels with a byte model) is the one timed in the
    # measured steps
    warm = {}
    if not args.no_roofline:
        ctx.kernel_timing(True)
    if args.warmup:
        scene.step(args.warmup)
    ctx.synchronize()
    if not args.no_roofline:
        ctx.kernel_timing(False)
        for nm in KERNELS:
            cnt, ms = ctx.kernel_time(nm)
            if cnt:
                warm[nm] = {"launches": cnt, "total_ms": round(ms, 4), "avg_us": round(1e3 * ms / cnt, 3)}
    print(f"[bench] warmup {args.warmup} done", file=sys.stderr, flush=True)

    n = spec.n
    dom = None
    if not args.no_roofline:
        dom = max((k for k in warm if k in BALG), key=lambda k: warm[k]["total_ms"], default="df_err_density")
        ctx.kernel_timing(True, only=[dom])
    s0 = scene.scalars()
    syncs = 0
    done = 0
    ctx.synchronize()
    t0 = time.perf_counter()
    while done < args.steps:           # progress lines every 500 steps
        k = min(500, args.steps - done)
        st = scene.step(k)
        syncs += st.host_syncs
        done += k
        if done < args.steps:
            print(f"[bench] {done} steps", file=sys.stderr, flush=True)
    ctx.synchronize()
    t1 = time.perf_counter()
    s1 = scene.scalars()
    wall = t1 - t0
    value = n * args.steps / wall
    it_d = (s1.sum_iterations - s0.sum_iterations) / args.steps
    it_v = (s1.sum_iterations_v - s0.sum_iterations_v) / args.steps
    mean_dt = (s1.time - s0.time) / args.steps
    print(f"[bench] timed {args.steps} steps: {wall * 1e3 / args.steps:.3f} ms/step, it_d {it_d:.2f} it_v {it_v:.2f}",
          file=sys.stderr, flush=True)
    step_bytes = balg_step(n, spec.nb, info["ncells"], it_d, it_v)

    roof = None
    kernels = {}
    if not args.no_roofline:
        ctx.kernel_timing(False)
        cnt, ms = ctx.kernel_time(dom)
        kernels[dom] = {"launches": cnt, "total_ms": round(ms, 4), "avg_us": round(1e3 * ms / cnt, 3) if cnt else 0.0}
        # Loop sweeps are launched ahead of the convergence read-back; a
        # launch for a converged scene returns at once.  The launches that do
        # the work are (1 + it_d) density error sweeps, it_d density velocity
        # sweeps and it_v of each divergence sweep per step; the returned
        # launches' time is charged to them.
        useful = {"df_err_density": args.steps * (1.0 + it_d), "df_vel_density": args.steps * it_d,
                  "df_err_divergence": args.steps * it_v, "df_vel_divergence": args.steps * it_v}
        per, bsw, csw = BALG[dom]
        bytes_launch = per * n + 16 * spec.nb * bsw + 4 * info["ncells"] * csw
        n_useful = useful.get(dom, kernels[dom]["launches"])
        avg_s = kernels[dom]["total_ms"] * 1e-3 / n_useful
        achieved = bytes_launch / avg_s / 1e9
        traffic = traffic_for(dom)
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom,
                # the measured HBM traffic per launch over the same duration
                "traffic_frac": round(traffic / avg_s / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
                "bytes_per_launch": bytes_launch, "avg_launch_us": kernels[dom]["avg_us"],
                "useful_launches": round(n_useful), "avg_useful_launch_us": round(avg_s * 1e6, 3)}
        print(f"[bench] kernels: {json.dumps(kernels)}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        import oracle.pyoracle as po
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
        x = scene.get(sphg.FIELD_X)
        v = scene.get(sphg.FIELD_V)
        k = scene.get(sphg.FIELD_K)
        kv = scene.get(sphg.FIELD_KV)
        sc = scene.scalars()
        o = po.Oracle(spec.params, spec.fluid, spec.boundary, fast=True, threads=threads)
        o.set_state(x, v, k, kv, sc.time_step, sc.iterations)
        c0 = time.perf_counter()
        o.step(args.cpu_steps)
        c1 = time.perf_counter()
        oit = o.scalars()
        cpu = {"value": round(n * args.cpu_steps / (c1 - c0), 1), "unit": "particle-steps/s", "cores": threads,
               "kind": "port",
               "sample": f"{args.cpu_steps} DFSPH steps of the same 1M-particle scene from the GPU's state after "
                         f"{args.warmup + args.steps} steps (last step: {int(oit['iterations'])} density iterations), "
                         "FP64 OpenMP restatement of the reference (-O3 -ffast-math -mavx2)"}
        o.close()

    scene.close()
    ctx.close()
    out = {
        "metric": "particle-steps/s (DFSPH 1M)",
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"DFSPH dam break, {n} fluid + {spec.nb} Akinci boundary particles, spacing 0.02, "
                        f"box 4x2.03x2.03, CFL 0.4, eta 0.01%, eta_V 0.1%, steps {args.warmup + 1}-"
                        f"{args.warmup + args.steps}",
            "storage": "f32 state, f64 neighbor sums",
            "n_fluid": n, "n_boundary": spec.nb, "cells": info["ncells"],
            "mean_density_iterations": round(it_d, 3), "mean_divergence_iterations": round(it_v, 3),
            "mean_dt": mean_dt, "host_syncs_per_step": round(syncs / max(1, args.steps), 3),
            "balg_bytes_per_step": int(step_bytes),
            "step_roofline_frac": round(step_bytes / (wall / args.steps) / (HBM_PEAK_GBS * 1e9), 4),
        },
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if kernels:
        out["kernels"] = kernels
    if warm:
        out["kernels_warmup"] = warm
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
