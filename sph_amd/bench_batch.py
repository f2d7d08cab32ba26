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


def compare(ctx, specs, steps: int, check: int = 8, log=None) -> dict:
    """Steps `specs` as one persistent batch, then each alone, `steps` steps
    each; returns both modes' aggregate fluid particle-steps/s and how many
    of the first `check` scenes ended bit-identical."""
    ntot = sum(s.n for s in specs)
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
    while done < steps:
        k = min(50, steps - done)
        st = b.run(k)
        syncs += st.host_syncs
        done += k
        if log:
            log(f"[batch] batched {done} steps")
    ctx.synchronize()
    t_b = time.perf_counter() - t0
    b.close()
    its = [s.scalars().sum_iterations for s in scenes]
    xb = [s.get(sphg.FIELD_X) for s in scenes[:check]]
    for s in scenes:
        s.close()
    # plain: each scene alone, created, stepped and freed in turn
    t_p = 0.0
    same = 0
    for i, spec in enumerate(specs):
        s = sphg.Scene(ctx, spec.params, spec.fluid, spec.boundary)
        ctx.synchronize()
        t0 = time.perf_counter()
        s.step(steps)
        ctx.synchronize()
        t_p += time.perf_counter() - t0
        if i < check:
            same += int(np.array_equal(s.get(sphg.FIELD_X), xb[i]))
        s.close()
        if log and (i + 1) % 32 == 0:
            log(f"[batch] plain {i + 1} scenes")
    work = float(ntot) * steps
    return {"scenes": len(specs), "fluid_particles": ntot, "steps": steps, "device_bytes": mem,
            "batched": {"particle_steps_per_s": round(work / t_b, 1), "seconds": round(t_b, 3),
                        "host_syncs_per_step": round(syncs / steps, 3)},
            "plain": {"particle_steps_per_s": round(work / t_p, 1), "seconds": round(t_p, 3)},
            "ratio": round(t_p / t_b, 3),
            "mean_density_iterations": round(float(np.mean(its)) / steps, 3),
            "setup_seconds": round(t_setup, 3),
            "identical_results": f"{same}/{min(check, len(specs))}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scenes", type=int, default=256)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--nmin", type=int, default=4096)
    ap.add_argument("--nmax", type=int, default=262144)
    ap.add_argument("--check", type=int, default=8, help="scenes whose batched and plain results are compared")
    args = ap.parse_args()

    specs = configs.batched_dfsph(args.scenes, args.nmin, args.nmax)
    print(f"[batch] {len(specs)} scenes, {sum(s.n for s in specs)} fluid particles", file=sys.stderr, flush=True)
    ctx = sphg.Context(0)
    r = compare(ctx, specs, args.steps, args.check, log=lambda m: print(m, file=sys.stderr, flush=True))
    ctx.close()
    out = {"metric": "aggregate particle-steps/s, batched DFSPH vs plain GPU solver",
           "workload": {"scenes": r["scenes"], "fluid_particles": r["fluid_particles"], "n_range": [args.nmin, args.nmax],
                        "steps": args.steps, "seed": 4, "device_bytes": r["device_bytes"]}}
    out.update({k: v for k, v in r.items() if k not in ("scenes", "fluid_particles", "steps", "device_bytes")})
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
