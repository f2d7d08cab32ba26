"""Layout tuner: block size and scan chunk as compiled variants.

The block size (threads per block = fluid particles per chunk) and the scan
chunk (cells per block of the counting sort's prefix scan) are compile-time
constants of the device layout: the chunk tables, the reduction partials and
the neighbor lists' groups are built on them, so a scene cannot switch them
at run time.  Each choice is therefore a library build of its own,

    make -C sph_amd/csrc BLOCK=128|256|512 CELL_CHUNK=<cells>

(sph_amd/libsphg_b<BLOCK>_c<CELL_CHUNK>.so; 256 x 1024 is libsphg.so), and
this tool times every built variant on a workload, each in a fresh process
(a process loads one libsphg), and names the fastest.  Per-particle results
do not depend on the layout; the solvers' convergence sums are added in
chunk order, so a variant may round a sum differently in its last FP64 bit,
and the tool reports each variant's density iterations and final-position
difference from the default build beside its time.

  python -m sph_amd.tune_layout [--workload dfsph_1m|batch64|wcsph_32k] [--build]

The reference has no counterpart: its loops are OpenMP `parallel for` over
particles (e.g. src/DFSPHSolver.cpp:256).
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT = (256, 1024)
VARIANTS = [(256, 1024), (128, 1024), (512, 1024), (256, 256), (256, 512), (256, 2048), (256, 4096)]
WORKLOADS = {
    # name: (warm-up steps, timed steps)
    "dfsph_1m": (10, 200),
    "batch64": (5, 30),
    "wcsph_32k": (20, 400),
    "dfsph_small": (5, 40),     # a 1,728-particle dam break (a quick check of identical results)
}


def lib_path(block: int, cells: int) -> str:
    if (block, cells) == DEFAULT:
        return os.path.join(HERE, "libsphg.so")
    return os.path.join(HERE, f"libsphg_b{block}_c{cells}.so")


def build(block: int, cells: int, jobs: int = 4) -> None:
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "csrc"), f"-j{jobs}", f"BLOCK={block}",
                    f"CELL_CHUNK={cells}"], check=True)


def _run_child(lib: str, workload: str, out_x: str) -> dict:
    """Times `workload` on library `lib` (in this process)."""
    from . import configs, sphg
    sphg.load(lib)
    warm, steps = WORKLOADS[workload]
    ctx = sphg.Context(0)
    if workload == "batch64":
        specs = configs.batched_dfsph(64)
        scenes = [sphg.Scene(ctx, s.params, s.fluid, s.boundary) for s in specs]
        b = sphg.Batch(ctx, scenes)
        b.run(warm)
        ctx.synchronize()
        t0 = time.perf_counter()
        b.run(steps)
        ctx.synchronize()
        wall = time.perf_counter() - t0
        b.close()
        work = sum(s.n for s in specs) * steps
        iters = sum(s.scalars().sum_iterations for s in scenes)
        x = scenes[0].get(sphg.FIELD_X)
        for s in scenes:
            s.close()
    else:
        spec = {"dfsph_1m": configs.dfsph_dam_1m, "wcsph_32k": configs.wcsph_dam_32k,
                "dfsph_small": lambda: configs.small_scene(sphg.DFSPH, n_side=12)}[workload]()
        s = sphg.Scene(ctx, spec.params, spec.fluid, spec.boundary)
        s.step(warm)
        ctx.synchronize()
        i0 = s.scalars().sum_iterations
        t0 = time.perf_counter()
        s.step(steps)
        ctx.synchronize()
        wall = time.perf_counter() - t0
        work = spec.n * steps
        iters = s.scalars().sum_iterations - i0
        x = s.get(sphg.FIELD_X)
        s.close()
    ctx.close()
    np.save(out_x, x)
    return {"ms_per_step": round(wall * 1e3 / steps, 4), "particle_steps_per_s": round(work / wall, 1),
            "density_iterations": int(iters)}


def tune(workload: str, variants, do_build: bool = False, timeout: float = 300.0, log=print) -> dict:
    """Times each variant in a child process; returns the table and the
    fastest variant."""
    rows = []
    tmp = os.environ.get("TMPDIR", "/tmp")
    ref_x = None
    for block, cells in variants:
        lib = lib_path(block, cells)
        if do_build:
            build(block, cells)
        if not os.path.exists(lib):
            log(f"[tune_layout] {os.path.basename(lib)} is not built; skipped")
            continue
        xf = os.path.join(tmp, f"tune_layout_{os.getpid()}_b{block}_c{cells}.npy")
        log(f"[tune_layout] {workload}: block {block}, scan chunk {cells}")
        p = subprocess.run([sys.executable, "-m", "sph_amd.tune_layout", "--child", lib, "--workload", workload,
                            "--out-x", xf], capture_output=True, text=True, timeout=timeout,
                           cwd=os.path.dirname(HERE))
        if p.returncode != 0:
            raise RuntimeError(f"variant b{block} c{cells} failed ({p.returncode}): {p.stderr[-2000:]}")
        r = json.loads(p.stdout.strip().splitlines()[-1])
        x = np.load(xf)
        os.remove(xf)
        if (block, cells) == DEFAULT:
            ref_x = x
        r.update({"block": block, "scan_chunk": cells})
        r["x_max_diff_vs_default"] = None if ref_x is None else float(np.abs(x - ref_x).max())
        rows.append(r)
        log(f"[tune_layout]   {r['ms_per_step']} ms/step, density iterations {r['density_iterations']}")
    best = min(rows, key=lambda r: r["ms_per_step"]) if rows else None
    return {"workload": workload, "variants": rows,
            "best": None if best is None else {"block": best["block"], "scan_chunk": best["scan_chunk"]}}


def main():
    ap = argparse.ArgumentParser(description="Time the block-size / scan-chunk library variants on a workload")
    ap.add_argument("--workload", default="dfsph_1m", choices=sorted(WORKLOADS))
    ap.add_argument("--variants", default=None, help="comma list of BLOCKxCELLS (default: all)")
    ap.add_argument("--build", action="store_true", help="build missing variants first")
    ap.add_argument("--child", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--out-x", default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.child:
        print(json.dumps(_run_child(args.child, args.workload, args.out_x)), flush=True)
        return
    variants = VARIANTS if not args.variants else [tuple(int(t) for t in v.split("x"))
                                                     for v in args.variants.split(",")]
    if variants[0] != DEFAULT and DEFAULT in variants:
        variants = [DEFAULT] + [v for v in variants if v != DEFAULT]
    print(json.dumps(tune(args.workload, variants, args.build, log=lambda s: print(s, flush=True))), flush=True)


if __name__ == "__main__":
    main()
