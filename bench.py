#!/usr/bin/env python3
"""Benchmark: DFSPH dam break, 1,000,000 FP32 fluid particles, one MI355X.

A step is one DFSPH solver step (src/DFSPHSolver.cpp:38-112 of the
reference) of the whole scene, inputs resident in HBM.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-extras]

Prints one JSON line.  `value` = fluid particle-steps/s over the K timed
steps after W untimed warm-up steps (host wall clock around the step loop,
bracketed by device synchronizes).  The workload's own run is 2000 steps:
the density iteration count per step changes over it (about 15 while the
column starts to fall, about 3 afterwards), so a short window is not
representative of it.  When K is below 2000 the line also carries
`config_run`: the whole 2000-step run, timed the same way in a scene of
its own.

The warm-up steps time every kernel with HIP event pairs on the library's
stream (`kernels_warmup`); the dominant one is then the only kernel an
event pair brackets during the timed steps, and only one launch in 4 of it
(an event pair costs the stream a few microseconds: timing every launch
took ~1.5% of the step).  `roofline.achieved` is that kernel's algorithmic
bytes per launch (per-particle figure x N plus boundary and cell terms)
divided by its average duration per launch that does work (solver loop
sweeps are also launched ahead for converged scenes, which return at once:
their time is charged to the working launches).  `roofline.traffic` is the
HBM traffic per launch of that kernel from the rocprofv3 FETCH_SIZE /
WRITE_SIZE summary in profiles/ (null when that summary has no figure for
the kernel), measured over the window `roofline.traffic_window` names.
`roofline.valu` is the same kernel's VALU ceiling: its VALU instructions per
launch from the rocprofv3 SQ_INSTS_VALU summary in profiles/, the time they
take to issue at MI355X's VALU rate (1024 SIMDs x 2.4 GHz, 4 cycles per
wave64 instruction: 78.6 TFLOP/s of FP64 FMA) and that time over this run's
working launch; `busy_measured` is SQ_ACTIVE_INST_VALU's own figure.

`cpu_baseline` times the FP64 OpenMP restatement of the reference
(oracle/liboracle_fast.so) on the host cores, one step from each of three
device states of the 2000-step run (its start, middle and end), with the
density iterations both solvers needed on those steps.

`batch` steps 64 scenes of the batched workload (N log-uniform in
4,096-262,144) for 100 steps as one batch and one at a time.  `serve`
serves a 24-request Poisson stream (WCSPH/PCISPH/DFSPH, N log-uniform in
2,048-2,000,000 with one request above 1M, 100-300 steps) with continuous
batching and one request at a time, with p50/p99 completion latency and the
throughput ratio the stream allows at best; `serve_burst` does the same for
64 requests of 2,048-262,144 particles arriving at once, and `serve_loaded`
for 100 requests of 2,048-2,000,000 particles arriving faster than the
device can serve them.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

TIMING_STRIDE = 4        # timed steps: one launch in 4 of the dominant kernel carries an event pair
HBM_PEAK_GBS = 8000.0
PMC_SUMMARY = os.path.join(ROOT, "profiles", "dfsph_1m_pmc.json")
SQ_SUMMARY = os.path.join(ROOT, "profiles", "dfsph_1m_sq_counters.json")
VALU_WAVE_INSTR_PER_S = 1024 * 2.4e9 / 4   # 1024 SIMDs at 2.4 GHz, a wave64 instruction every 4 cycles
CONFIG_STEPS = 2000

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


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def balg_step(n, nb, ncells, it_d, it_v):
    """Algorithmic HBM bytes of one DFSPH step: 264 B per particle fixed,
    88 B per particle per density and per divergence iteration, 16 B per
    boundary particle for each of the 4 + 2 it_d + 2 it_v boundary sweeps,
    4 B per cell per neighbor sweep and 12 B per cell per grid build."""
    sweeps = 4 + 2 * it_d + 2 * it_v
    return (264 + 88 * (it_d + it_v)) * n + 16 * nb * sweeps + 4 * ncells * sweeps + 12 * ncells


def pmc_summary(path=PMC_SUMMARY):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def valu_ceiling(kernel, avg_s):
    """VALU issue time of one launch of `kernel` (profiles/ SQ summary) and
    its fraction of the measured working launch `avg_s`; None without a
    figure for the kernel."""
    sq = pmc_summary(SQ_SUMMARY)
    k = sq.get("kernels", {}).get(kernel)
    if not k:
        return None
    instr = k["valu_per_wavefront"] * k["wavefronts"]
    issue_s = instr / VALU_WAVE_INSTR_PER_S
    return {"instr_per_launch": round(instr), "issue_us": round(issue_s * 1e6, 2),
            "peak_wave_instr_per_s": VALU_WAVE_INSTR_PER_S, "frac": round(issue_s / avg_s, 4),
            "busy_measured": k.get("valu_busy"), "window": sq.get("workload")}


def timed_steps(scene, ctx, steps, capture=(), state_of=None):
    """Steps the scene `steps` times on the wall clock (chunks of at most 500
    steps, a progress line after each); at the step counts in `capture` the
    clock is paused and state_of(scene) is kept.  Returns (seconds, host
    syncs, {step: state})."""
    done, syncs, wall, states = 0, 0, 0.0, {}
    marks = sorted(set(c for c in capture if 0 <= c < steps))
    while done < steps:
        at_mark = bool(marks) and marks[0] == done
        if at_mark:
            states[done] = state_of(scene)
            marks.pop(0)
        # a captured state's next step runs alone, so that its iteration
        # count can be read after it
        stop = done + 1 if at_mark else min([done + 500, steps] + [m for m in marks if m > done])
        ctx.synchronize()
        t0 = time.perf_counter()
        st = scene.step(stop - done)
        ctx.synchronize()
        wall += time.perf_counter() - t0
        syncs += st.host_syncs
        if at_mark:
            states[done]["gpu_it"] = scene.scalars().iterations
        done = stop
        if done < steps and done % 500 == 0:
            log(f"{done} steps")
    return wall, syncs, states


def device_state(scene):
    from sph_amd import sphg
    sc = scene.scalars()
    return {"x": scene.get(sphg.FIELD_X), "v": scene.get(sphg.FIELD_V), "k": scene.get(sphg.FIELD_K),
            "kv": scene.get(sphg.FIELD_KV), "dt": sc.time_step, "it": sc.iterations,
            "sum_it": sc.sum_iterations, "step": sc.steps}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=CONFIG_STEPS)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n-side", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="no config_run, batch or serve objects")
    ap.add_argument("--no-roofline", action="store_true", help="no per-kernel events in the timed steps")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the solver has no multi-GPU path (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)

    from sph_amd import configs, sphg

    spec = configs.dfsph_dam_1m(n_side=args.n_side)
    log(f"scene {spec.name}: {spec.n} fluid, {spec.boundary.shape[0]} boundary")
    ctx = sphg.Context(0)
    scene = sphg.Scene(ctx, spec.params, spec.fluid, spec.boundary)
    info = scene.info()
    n = spec.n

    # ---------------------------------------------------------- warm-up
    # every kernel timed; the dominant one (largest total time among the
    # kernels with a byte model) is the one timed in the measured steps
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
    log(f"warmup {args.warmup} done")

    # ------------------------------------------------------ timed steps
    dom = None
    if not args.no_roofline:
        dom = max((k for k in warm if k in BALG), key=lambda k: warm[k]["total_ms"], default="df_err_density")
        ctx.kernel_timing(True, only=[dom])
        # every TIMING_STRIDE-th launch carries an event pair: a sample of
        # the launches at a quarter of the pairs' cost to the timed steps
        ctx.kernel_timing_stride(TIMING_STRIDE)
    whole = args.warmup + args.steps >= CONFIG_STEPS
    samples = (0, args.steps // 2, args.steps - 1) if whole and not args.no_cpu else ()
    s0 = scene.scalars()
    wall, syncs, states = timed_steps(scene, ctx, args.steps, samples, device_state)
    s1 = scene.scalars()
    value = n * args.steps / wall
    it_d = (s1.sum_iterations - s0.sum_iterations) / args.steps
    it_v = (s1.sum_iterations_v - s0.sum_iterations_v) / args.steps
    mean_dt = (s1.time - s0.time) / args.steps
    log(f"timed {args.steps} steps: {wall * 1e3 / args.steps:.3f} ms/step, it_d {it_d:.2f} it_v {it_v:.2f}")
    step_bytes = balg_step(n, spec.nb, info["ncells"], it_d, it_v)

    roof = None
    kernels = {}
    if not args.no_roofline:
        ctx.kernel_timing(False)
        cnt, ms = ctx.kernel_time(dom)
        # the sampled launches stand for TIMING_STRIDE launches each
        kernels[dom] = {"launches": cnt * TIMING_STRIDE, "total_ms": round(ms * TIMING_STRIDE, 4),
                        "avg_us": round(1e3 * ms / cnt, 3) if cnt else 0.0, "sampled_launches": cnt,
                        "sample_stride": TIMING_STRIDE}
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
        pmc = pmc_summary()
        k = pmc.get("kernels", {}).get(dom)
        traffic = int(k["fetch_bytes"] + k["write_bytes"]) if k else None
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom,
                "traffic_window": pmc.get("workload") if traffic else None,
                # the measured HBM traffic per launch (PMC window) over this
                # run's launch duration
                "traffic_frac": round(traffic / avg_s / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
                "bytes_per_launch": bytes_launch, "avg_launch_us": kernels[dom]["avg_us"],
                "useful_launches": round(n_useful), "avg_useful_launch_us": round(avg_s * 1e6, 3),
                "valu": valu_ceiling(dom, avg_s)}
        log(f"kernels: {json.dumps(kernels)}")
    scene.close()

    # -------------------------------------------- the workload's own run
    config_run = None
    if not args.no_extras and not whole:
        sc2 = sphg.Scene(ctx, spec.params, spec.fluid, spec.boundary)
        q0 = sc2.scalars()
        cs = (0, CONFIG_STEPS // 2, CONFIG_STEPS - 1) if not args.no_cpu else ()
        w2, sy2, states = timed_steps(sc2, ctx, CONFIG_STEPS, cs, device_state)
        q1 = sc2.scalars()
        cit_d = (q1.sum_iterations - q0.sum_iterations) / CONFIG_STEPS
        cit_v = (q1.sum_iterations_v - q0.sum_iterations_v) / CONFIG_STEPS
        cb = balg_step(n, spec.nb, info["ncells"], cit_d, cit_v)
        config_run = {"steps": CONFIG_STEPS, "value": round(n * CONFIG_STEPS / w2, 1),
                      "ms_per_step": round(w2 * 1e3 / CONFIG_STEPS, 4), "mean_density_iterations": round(cit_d, 3),
                      "mean_divergence_iterations": round(cit_v, 3),
                      "host_syncs_per_step": round(sy2 / CONFIG_STEPS, 3),
                      "step_roofline_frac": round(cb / (w2 / CONFIG_STEPS) / (HBM_PEAK_GBS * 1e9), 4)}
        log(f"config run: {config_run['ms_per_step']} ms/step, it_d {cit_d:.2f}")
        sc2.close()

    # ------------------------------------------------------ CPU baseline
    cpu = None
    if not args.no_cpu and states:
        import oracle.pyoracle as po
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
        o = po.Oracle(spec.params, spec.fluid, spec.boundary, fast=True, threads=threads)
        cpu_s, rows = 0.0, []
        for step in sorted(states):
            st = states[step]
            o.set_state(st["x"], st["v"], st["k"], st["kv"], st["dt"], st["it"])
            c0 = time.perf_counter()
            o.step(1)
            cpu_s += time.perf_counter() - c0
            rows.append((int(st["step"]) + 1, int(o.scalars()["iterations"]), int(st["gpu_it"])))
        o.close()
        cpu = {"value": round(n * len(rows) / cpu_s, 1), "unit": "particle-steps/s", "cores": threads,
               "kind": "port",
               "sample": "one DFSPH step of the same 1M-particle scene from each of the device states before steps "
                         + ", ".join(str(r[0]) for r in rows)
                         + f" of its {CONFIG_STEPS}-step run (density iterations on those steps: CPU "
                         + ", ".join(str(r[1]) for r in rows) + ", GPU " + ", ".join(str(r[2]) for r in rows)
                         + "); FP64 OpenMP restatement of the reference (-O3 -ffast-math -mavx2)",
               "density_iterations": [r[1] for r in rows], "gpu_density_iterations": [r[2] for r in rows],
               "gpu_value_same_steps_note": "the GPU value is over the whole timed window; its mean density "
                                            "iterations per step are in config / config_run"}
        log(f"cpu baseline {cpu['value']:.3g} particle-steps/s on {threads} threads")

    # ---------------------------------------------- batching and serving
    batch = serve = burst = loaded = None
    if not args.no_extras:
        from sph_amd import bench_batch, serve as srv
        bspecs = configs.batched_dfsph(64)
        batch = bench_batch.compare(ctx, bspecs, 100, check=8)
        batch["workload"] = "64 scenes of the batched DFSPH workload (N log-uniform 4,096-262,144, seed 4), 100 steps"
        log(f"batch: ratio {batch['ratio']}")
        seed = 5
        while True:
            reqs = srv.request_stream(24, 30.0, 2048, 2_000_000, 100, 300, seed)
            if max(q.n_target for q in reqs) >= 1_000_000:
                break
            seed += 1
        serve = srv.compare(ctx, reqs, 200 * 2 ** 30, quantum=25, check=4, parallel=True, split_n=srv.SPLIT_N)
        serve["stream"] = (f"24 requests, Poisson 30/s, WCSPH/PCISPH/DFSPH uniform, N log-uniform 2,048-2,000,000 "
                           f"(largest {max(q.n_target for q in reqs)}), 100-300 steps, seed {seed}, 200 GB budget, "
                           f"quantum 25, lanes by method and size (split at {srv.SPLIT_N}): each small lane on a "
                           f"stream of its own, the large lanes on one stream, the oldest request's lane first, "
                           f"stepping its oldest scenes up to {srv.LARGE_CAP} fluid particles")
        log(f"serve: ratio {serve['throughput_ratio']}, p99 {serve['batched']['latency_p99_s']} s")
        reqs = srv.request_stream(64, 1000.0, 2048, 262144, 100, 300, 5)
        burst = srv.compare(ctx, reqs, 200 * 2 ** 30, quantum=25, check=4, parallel=True, split_n=srv.SPLIT_N)
        burst["stream"] = ("64 requests, Poisson 1000/s (a burst), WCSPH/PCISPH/DFSPH uniform, N log-uniform "
                           "2,048-262,144, 100-300 steps, seed 5, 200 GB budget, quantum 25, lanes by method and "
                           f"size (split at {srv.SPLIT_N}) as in `serve`")
        log(f"serve burst: ratio {burst['throughput_ratio']}, p99 {burst['batched']['latency_p99_s']} s")
        reqs = srv.request_stream(100, 30.0, 2048, 2_000_000, 100, 300, 11)
        loaded = srv.compare(ctx, reqs, 200 * 2 ** 30, quantum=25, check=4, parallel=True, split_n=srv.SPLIT_N)
        loaded["stream"] = ("100 requests, Poisson 30/s (a queue builds up: ~3.3 s of arrivals for ~9 s of work), "
                            "WCSPH/PCISPH/DFSPH uniform, N log-uniform 2,048-2,000,000, 100-300 steps, seed 11, "
                            "200 GB budget, quantum 25, lanes as in `serve`")
        log(f"serve loaded: ratio {loaded['throughput_ratio']}, p50 {loaded['batched']['latency_p50_s']} s, "
            f"p99 {loaded['batched']['latency_p99_s']} s")

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
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"DFSPH dam break, {n} fluid + {spec.nb} Akinci boundary particles, spacing 0.02, "
                        f"box 4x2.03x2.03, CFL 0.4, eta 0.01%, eta_V 0.1%, steps {args.warmup + 1}-"
                        f"{args.warmup + args.steps}",
            "storage": "f32 state; neighbor sums and the solver's scalars in f64",
            "n_fluid": n, "n_boundary": spec.nb, "cells": info["ncells"],
            "mean_density_iterations": round(it_d, 3), "mean_divergence_iterations": round(it_v, 3),
            "mean_dt": mean_dt, "host_syncs_per_step": round(syncs / max(1, args.steps), 3),
            "balg_bytes_per_step": int(step_bytes),
            "step_roofline_frac": round(step_bytes / (wall / args.steps) / (HBM_PEAK_GBS * 1e9), 4),
        },
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if config_run:
        out["config_run"] = config_run
    if batch:
        out["batch"] = batch
    if serve:
        out["serve"] = serve
    if burst:
        out["serve_burst"] = burst
    if loaded:
        out["serve_loaded"] = loaded
    if kernels:
        out["kernels"] = kernels
    if warm:
        out["kernels_warmup"] = warm
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
