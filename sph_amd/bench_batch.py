"""Batched DFSPH benchmark: many independent scenes, one launch per solver
phase, against the plain one-scene-at-a-time GPU solver on the same GPU.

Workload (defaults): 256 DFSPH dam-break scenes, fluid sizes log-uniform in
[4,096, 262,144] (seeded), each a jittered lattice block filling 20-60 % of
its box, spacing 0.02, per-scene adaptive dt, tolerances of the 1M dam
break, 500 steps per scene.  Prints one JSON line with the aggregate fluid
particle-steps/s of both modes.

  python -m sph_amd.bench_batch [--scenes 256] [--steps 500] [--sample 0]
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import configs, sphg


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scenes", type=int, default=256)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--nmin", type=int, default=4096)
    ap.add_argument("--nmax", type=int, default=262144)
    ap.add_argument("--check", type=int, default=8, help="scenes whose batched and plain results are compared")
    args = ap.parse_args()

    specs = configs.batched_dfsph(args.scenes, args.nmin, args.nmax)
    ntot = sum(s.n for s in specs)
    print(f"[batch] {len(specs)} scenes, {ntot} fluid particles", file=sys.stderr, flush=True)
    ctx = sphg.Context(0)

    # batched: every scene resident, one launch per phase for all of them
    t0 = time.perf_counter()
    scenes = [sphg.Scene(ctx, s.params, s.fluid, s.boundary) for s in specs]
    ctx.synchronize()
    t_setup = time.perf_counter() - t0
    mem = sum(s.info()["device_bytes"] for s in scenes)
    b = sphg.Batch(ctx, scenes)
    ctx.synchronize()
    t0 = time.perf_counter()
    done = 0
    syncs = 0
    while done < args.steps:
        k = min(50, args.steps - done)
        st = b.run(k)
        syncs += st.host_syncs
        done += k
        print(f"[batch] batched {done} steps", file=sys.stderr, flush=True)
    ctx.synchronize()
    t_b = time.perf_counter() - t0
    b.close()
    its = [s.scalars().sum_iterations for s in scenes]
    xb = [s.get(sphg.FIELD_X) for s in scenes[: args.check]]
    for s in scenes:
        s.close()

    # plain: each scene alone, created, stepped and freed in turn
    t_p = 0.0
    same = 0
    for i, spec in enumerate(specs):
        s = sphg.Scene(ctx, spec.params, spec.fluid, spec.boundary)
        ctx.synchronize()
        t0 = time.perf_counter()
        s.step(args.steps)
        ctx.synchronize()
        t_p += time.perf_counter() - t0
        if i < args.check:
            same += int(np.array_equal(s.get(sphg.FIELD_X), xb[i]))
        s.close()
        if (i + 1) % 32 == 0:
            print(f"[batch] plain {i + 1} scenes", file=sys.stderr, flush=True)
    ctx.close()
    work = float(ntot) * args.steps
    out = {"metric": "aggregate particle-steps/s, batched DFSPH vs plain GPU solver",
           "workload": {"scenes": len(specs), "fluid_particles": ntot, "n_range": [args.nmin, args.nmax],
                        "steps": args.steps, "seed": 4, "device_bytes": mem},
           "batched": {"particle_steps_per_s": round(work / t_b, 1), "seconds": round(t_b, 3),
                       "host_syncs_per_step": round(syncs / args.steps, 3)},
           "plain": {"particle_steps_per_s": round(work / t_p, 1), "seconds": round(t_p, 3)},
           "ratio": round(t_p / t_b, 3),
           "mean_density_iterations": round(float(np.mean(its)) / args.steps, 3),
           "setup_seconds": round(t_setup, 3),
           "identical_results": f"{same}/{min(args.check, len(specs))}"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
