"""Serving: simulation requests on one GPU, continuous batching vs plain.

The reference is a one-scene program (its Simulation singleton,
src/Simulation.cpp:84-90); this module has no counterpart there.  A request
asks for `steps` solver steps of a seeded scene (method, fluid size,
geometry); its result is the fluid state after those steps, on the host.

Continuous batching (`serve_batched`):
  * requests arrive on a wall clock (Poisson process); a waiting request is
    admitted, first come first served, when the device bytes of its scene
    (known before allocation, `Scene.plan_bytes`) fit the memory budget
    beside the scenes already resident; its scene is built, created and
    planned on a geometry thread on its lane group's context (host-side
    calls the library allows beside a running context), and initialised
    there by the group's own thread before the quantum that first steps it;
  * resident scenes stay on the device, grouped in lanes: one per solver
    method, split by size (`split_n`) so that small scenes do not wait on
    the step time of giant ones; a lane's scenes are stepped together, one
    launch per solver phase covering all of them (`sphg.Batch`), in quanta
    of at most `quantum` steps, so that requests that arrive meanwhile join
    at the lane's next quantum;
  * a scene that has done its steps is read back, freed and its request
    completed; the rest continue;
  * a step call that reports a neighbor overflow (SPHG_EOVERFLOW: some
    particle had more neighbors than its list holds) fails only the requests
    whose scenes overflowed: they are freed and completed with an error, the
    batch is rebuilt without them, and the other scenes, which stepped
    normally, continue;
  * lanes run on the context (HIP stream) of their group (`group_of`):
    each small lane has its own, and the large lanes share one, four streams
    for the process's four hardware queues; a group runs one quantum at a
    time, the lane with the oldest request first, and a large lane steps its
    oldest scenes up to `large_cap` fluid particles (a large scene fills the
    device alone: stepping large requests in arrival order finishes each
    sooner, at the same throughput);
  * with `parallel`, the groups run their quanta at the same time and
    independently, each driven by its own host thread: a launch-bound lane
    of small scenes then overlaps the large scenes instead of waiting for
    them, and a group whose quantum ends is refilled at once.
  The scheduler thread only admits, launches and retires: the scene set-up
  and read-back calls run on the geometry threads and the groups' threads.

Plain serving (`serve_plain`): the same requests one at a time in arrival
order, each stepped alone from creation to read-back.

Both report per-request completion latency (arrival to result on the host)
and the aggregate fluid particle-steps/s over the stream.

  python -m sph_amd.serve [--requests 48] [--rate 2.0] [--nmax 262144] ...
"""
from __future__ import annotations

import argparse
import json
import math
import threading
import time
from collections import deque
from concurrent.futures import FIRST_COMPLETED, Future, ThreadPoolExecutor, wait
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional

import numpy as np

from . import configs, sphg
from .configs import SceneSpec, SplitMix64
from .sphg import DFSPH, PCISPH, WCSPH, SceneParams

METHODS = (WCSPH, PCISPH, DFSPH)
METHOD_NAMES = {WCSPH: "WCSPH", PCISPH: "PCISPH", DFSPH: "DFSPH"}


@dataclass
class Request:
    rid: int
    arrival: float      # seconds after the stream starts
    method: int
    n_target: int
    steps: int
    geometry: str       # "block" (dam break) or "drop" (sphere drop)
    seed: int


def request_stream(count: int, rate: float, nmin: int = 2048, nmax: int = 2_000_000, smin: int = 100,
                   smax: int = 2000, seed: int = 5) -> List[Request]:
    """Poisson arrivals at `rate` per second (dt = -ln(1 - u) / rate); method
    uniform over WCSPH, PCISPH, DFSPH; fluid size log-uniform in [nmin, nmax];
    steps uniform in [smin, smax]; geometry block or sphere drop."""
    rng = SplitMix64(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += -math.log(1.0 - rng.uniform()) / rate
        method = METHODS[min(2, int(rng.uniform() * 3))]
        n = rng.log_uniform_int(nmin, nmax)
        steps = smin + min(smax - smin, int(rng.uniform() * (smax - smin + 1)))
        geometry = "block" if rng.uniform() < 0.5 else "drop"
        out.append(Request(i, t, method, n, steps, geometry, 10_000 + i))
    return out


def _method_params(method: int, r: float) -> SceneParams:
    p = SceneParams(method=method, particle_radius=r, gravity=configs.G)
    if method == DFSPH:
        return SceneParams(method=DFSPH, particle_radius=r, gravity=configs.G, cfl_factor=0.4, time_step=1e-3,
                           max_time_step=5e-3, min_time_step=1e-6, eta=1e-4, eta_v=1e-3, min_iterations=2,
                           max_iterations=100, min_iterations_v=1, max_iterations_v=100)
    if method == WCSPH:
        return SceneParams(method=WCSPH, particle_radius=r, gravity=configs.G, stiffness=1000.0, gamma=7.0,
                           viscosity_method=sphg.VISCOSITY_XSPH, viscosity=0.01, time_step=5e-4,
                           min_time_step=5e-4, max_time_step=5e-4)
    if method == PCISPH:
        return SceneParams(method=PCISPH, particle_radius=r, gravity=configs.G, eta=0.01, min_iterations=3,
                           max_iterations=100, time_step=5e-4, min_time_step=5e-4, max_time_step=5e-4)
    return p


def build_spec(req: Request, spacing: float = 0.02) -> SceneSpec:
    """The seeded scene of a request: a jittered lattice block in the corner
    of a box twice its length ("block"), or a lattice sphere released half a
    diameter above the floor of a box ("drop"), Akinci walls, spacing 0.02."""
    r = spacing / 2.0
    h = 4.0 * r
    params = _method_params(req.method, r)
    if req.geometry == "block":
        ny = max(2, int(round((req.n_target / 2.0) ** (1.0 / 3.0))))
        nx = max(2, int(round(req.n_target / (ny * ny))))
        L = (nx, ny, ny)
        box = (2.0 * nx * spacing + 6 * r, ny * spacing + 6 * r, 1.5 * ny * spacing + 6 * r)
        spec = configs.dam_break(req.method, L, spacing, box, jitter=0.01, seed=req.seed, gap=2.5)
        spec.params = params
    else:
        # sphere of ~n_target lattice sites: radius from n s^3 = 4/3 pi R^3
        R = (3.0 * req.n_target / (4.0 * math.pi)) ** (1.0 / 3.0) * spacing
        side = 2.0 * R + 12 * r
        c = (side / 2.0, side / 2.0, 2.0 * R + 6 * r)
        box = (side, side, c[2] + R + 6 * r)
        fluid = sphg.fluid_spheres((0.0, 0.0, 0.0), box, r, [c], [R])
        boundary = sphg.sample_box((0.0, 0.0, 0.0), box, h)
        spec = SceneSpec("drop", params, fluid, boundary)
    spec.name = f"req{req.rid}_{METHOD_NAMES[req.method]}_{req.geometry}_{spec.n}"
    return spec


# --------------------------------------------------------------- backends

SMALL, LARGE = 0, 1
SPLIT_N = 262144     # default size class boundary, fluid particles
LARGE_CAP = 1_000_000   # default: a large-scene lane steps its oldest scenes up to 1M fluid particles


def lane_of(method: int, n: int, split_n: int = 0):
    """The batch a resident scene steps in: its method and, when `split_n` is
    positive, its size class (LARGE from `split_n` fluid particles up).  One
    batch step lasts as long as its biggest scene's step, so a small scene
    that shares a batch with a 2M-particle one advances at the big scene's
    pace; in a lane of its own it advances at its own."""
    return (method, LARGE if split_n and n >= split_n else SMALL)


def group_of(lane, split_n: int = 0):
    """The context (one HIP stream) a lane's batches run on.  With size
    classes, the large lanes of all methods share one and take turns (a
    large scene fills the device on its own), and every small lane has its
    own: four streams for three methods, one per hardware queue of the
    process's default four.  Streams beyond that share hardware queues, and
    a small scene's short kernels then wait in order behind a large scene's
    on its queue.  Without size classes every lane has its own."""
    return (None, LARGE) if split_n and lane[1] == LARGE else lane


class GpuBackend:
    """The HIP solver library on one device.  With `lane_contexts` each lane
    group (`group_of`) gets a context (HIP stream) of its own, so that the
    batches of different groups can run concurrently.
    The scheduler then creates and plans a lane's new scenes (host-side calls
    the library allows beside a running context) while the lane's quantum
    runs, and initialises, batches and closes them only while it is idle."""

    def __init__(self, ctx: sphg.Context, lane_contexts: bool = False):
        self.ctx = ctx
        self.lane_contexts = lane_contexts
        self.extra = []
        self.ctxs = {}
        self.lock = threading.Lock()    # scenes are created on the geometry threads

    def _context(self, lane):
        if lane is None or not self.lane_contexts:
            return self.ctx
        with self.lock:
            c = self.ctxs.get(lane)
            if c is None:
                if self.ctxs:
                    c = sphg.Context(self.ctx.device)
                    self.extra.append(c)
                else:
                    c = self.ctx
                self.ctxs[lane] = c
            return c

    def create(self, spec: SceneSpec, lane=None):
        return sphg.Scene(self._context(lane), spec.params, spec.fluid, spec.boundary, init=False)

    def plan_bytes(self, scene) -> int:
        return scene.plan_bytes()

    def init(self, scene) -> None:
        scene.init()

    def batch(self, scenes):
        return sphg.Batch(scenes[0].ctx, scenes)

    def step(self, scene, k: int) -> None:
        scene.step(k)

    def result(self, scene) -> np.ndarray:
        return scene.get(sphg.FIELD_X)

    def overflowed(self, scene) -> bool:
        """Whether the scene's last step call had a neighbor overflow."""
        return scene.scalars().neighbor_overflow > 0

    def close(self, obj) -> None:
        obj.close()

    def shutdown(self) -> None:
        """Closes the contexts this backend created."""
        for c in self.extra:
            c.close()
        self.extra = []
        self.ctxs = {}


@dataclass
class Job:
    req: Request
    spec_future: object = None
    scene: object = None
    lane: tuple = None
    group: tuple = None
    bytes: int = 0
    n: int = 0
    remaining: int = 0
    admitted: float = -1.0
    inited: bool = False
    done: float = -1.0
    result: Optional[np.ndarray] = None
    error: Optional[str] = None      # set when the request failed (completed without a result)


@dataclass
class ServeResult:
    mode: str
    jobs: List[Job]
    makespan: float
    rejected: List[int] = field(default_factory=list)

    def summary(self) -> dict:
        done = [j for j in self.jobs if j.done >= 0 and j.error is None]
        failed = [j for j in self.jobs if j.error is not None]
        lat = np.array([j.done - j.req.arrival for j in done]) if done else np.zeros(1)
        work = float(sum(j.n * j.req.steps for j in done))
        return {"mode": self.mode, "requests": len(done), "rejected": len(self.rejected), "failed": len(failed),
                "particle_steps": work, "makespan_s": round(self.makespan, 3),
                "particle_steps_per_s": round(work / self.makespan, 1) if self.makespan > 0 else None,
                "latency_p50_s": round(float(np.percentile(lat, 50)), 4),
                "latency_p99_s": round(float(np.percentile(lat, 99)), 4),
                "latency_max_s": round(float(lat.max()), 4)}


def _spec_pool():
    return ThreadPoolExecutor(max_workers=2, thread_name_prefix="spec")


def _is_overflow(e: BaseException) -> bool:
    return isinstance(e, sphg.SphgError) and e.code == sphg.EOVERFLOW


def _run_quantum(batch, k):
    """Runs a batch quantum; returns the neighbor-overflow error it raised,
    if any (the other scenes of the batch have stepped normally), and lets
    every other error propagate."""
    try:
        batch.run(k)
    except sphg.SphgError as e:
        if not _is_overflow(e):
            raise
        return e
    return None


def _prepare_and_run(backend, fresh, old_batch, scenes, batch, k, finishing=()):
    """One quantum of a lane on its group's thread: initialises the scenes
    stepping for the first time, builds the batch when its scenes changed
    (closing the one it replaces), runs the quantum and reads back the
    scenes it finishes.  Returns the batch, the overflow error
    `_run_quantum` reports and the read-backs (None for a scene that
    overflowed)."""
    if batch is None and old_batch is not None:
        backend.close(old_batch)        # first: no other reference to it is kept
    for sc in fresh:
        backend.init(sc)
    if batch is None:
        batch = backend.batch(scenes)
    try:
        over = _run_quantum(batch, k)
    except BaseException:
        backend.close(batch)            # a batch is closed before its scenes
        raise
    results = [None if over is not None and backend.overflowed(sc) else backend.result(sc) for sc in finishing]
    return batch, over, results


@dataclass
class _Lane:
    group: tuple = None                             # its context (`group_of`)
    jobs: List[Job] = field(default_factory=list)   # resident scenes, admission order
    batch: object = None                            # sphg.Batch over `stepping`, or None
    stepping: List[Job] = field(default_factory=list)   # the jobs `batch` steps
    flight: object = None                           # (future, jobs, k) of the running quantum
    turn: int = 0                                   # launches so far of all lanes when this one last launched


def serve_batched(backend, requests: List[Request], budget_bytes: int, quantum: int = 25,
                  clock: Callable[[], float] = time.perf_counter, sleep: Callable[[float], None] = time.sleep,
                  keep_results: bool = True, log: Callable[[str], None] = None,
                  parallel: bool = False, split_n: int = 0, large_cap: int = 0) -> ServeResult:
    """Continuous batching of `requests` (sorted by arrival) on `backend`.

    Resident scenes are grouped in lanes (`lane_of`: method, and size class
    when `split_n` > 0); each lane's scenes step together as one batch.
    Lanes run on the context of their group (`group_of`); the lanes of one
    group take turns.  With `parallel` every group runs its quanta on a host
    thread of its own (the backend gives each group its own stream) and
    independently of the others: when a quantum ends, its lane's finished
    scenes are completed, arrivals admitted and the group's next quantum
    started while the other groups run on.  Without it the lanes take turns
    on the calling thread.  With `large_cap` a large lane steps only its
    oldest scenes, up to `large_cap` fluid particles (at least one)."""
    if parallel and not getattr(backend, "lane_contexts", True):
        raise ValueError("parallel lanes need a backend with a context per lane group "
                         "(GpuBackend(lane_contexts=True)): a context is driven by one host thread at a time")
    pending = deque(sorted(requests, key=lambda q: q.arrival))
    waiting: deque = deque()
    lanes: Dict[tuple, _Lane] = {}
    jobs: List[Job] = []
    rejected: List[int] = []
    used = 0
    launches = 0
    pool = _spec_pool()
    runner = ThreadPoolExecutor(max_workers=2 * len(METHODS), thread_name_prefix="lane") if parallel else None
    t0 = clock()

    def prepare(req):
        """On a geometry thread: the request's scene, created and planned on
        its lane group's context (host-side calls, allowed beside the
        context's running quanta)."""
        spec = build_spec(req)
        lane = lane_of(req.method, spec.n, split_n)
        group = group_of(lane, split_n)
        scene = backend.create(spec, group)
        try:
            return spec.n, lane, group, scene, backend.plan_bytes(scene)
        except BaseException:
            backend.close(scene)
            raise

    def admit(block: bool):
        """Admits waiting requests in arrival order while they fit.  Unless
        `block`, it stops at a request whose scene is still being prepared
        (the lanes keep running meanwhile)."""
        nonlocal used
        while waiting:
            j = waiting[0]
            if j.scene is None:
                if not block and not j.spec_future.done():
                    break
                j.n, j.lane, j.group, j.scene, j.bytes = j.spec_future.result()
            if j.bytes > budget_bytes:
                backend.close(j.scene)
                j.scene = None
                rejected.append(j.req.rid)
                waiting.popleft()
                continue
            if used + j.bytes > budget_bytes:
                break
            # its device memory is reserved now; the scene is initialised
            # (allocated and uploaded) on its group's context when its lane
            # next steps it, while no quantum of that group is running
            used += j.bytes
            j.remaining = j.req.steps
            j.admitted = clock() - t0
            ln = lanes.setdefault(j.lane, _Lane(group=j.group))
            ln.jobs.append(j)
            waiting.popleft()

    def launch(ln: _Lane):
        # a large lane steps its oldest scenes up to `large_cap` fluid
        # particles (a large scene fills the device on its own: batching more
        # delays the first, and their lists outgrow the Infinity Cache); the
        # others wait, resident, for their turn
        js = list(ln.jobs)
        if large_cap and ln.jobs[0].lane[1] == LARGE:
            # oldest first, up to large_cap fluid particles (at least one scene)
            tot, k = 0, 0
            while k < len(js) and (k == 0 or tot + js[k].n <= large_cap):
                tot += js[k].n
                k += 1
            js = js[:k]
        # the group's thread initialises the new scenes and rebuilds the
        # batch before the quantum, so this thread goes on serving the other
        # groups meanwhile
        fresh = [j for j in js if not j.inited]
        for j in fresh:
            j.inited = True
        keep = ln.batch is not None and [id(j) for j in js] == [id(j) for j in ln.stepping]
        k = min(quantum, min(j.remaining for j in js))
        args = (backend, [j.scene for j in fresh], None if keep else ln.batch, [j.scene for j in js],
                ln.batch if keep else None, k, [j.scene for j in js if j.remaining == k])
        ln.batch, ln.stepping = None, js    # the quantum's batch comes back with it
        if runner is not None:
            fut = runner.submit(_prepare_and_run, *args)
        else:
            fut = Future()
            try:
                fut.set_result(_prepare_and_run(*args))
            except BaseException as e:  # noqa: BLE001 - re-raised by land()
                fut.set_exception(e)
        ln.flight = (fut, js, k)

    def land(ln: _Lane):
        """The lane's quantum has ended: account its steps, fail overflowed
        scenes, complete finished ones."""
        nonlocal used
        fut, js, k = ln.flight
        ln.flight = None
        # raises what the quantum raised, other than an overflow
        ln.batch, over, results = fut.result()
        read = dict(zip([id(j) for j in js if j.remaining == k], results))
        if over is not None:
            bad = [j for j in js if backend.overflowed(j.scene)]
            if not bad:
                raise over
            backend.close(ln.batch)
            ln.batch = None
            for j in bad:
                ln.jobs.remove(j)
                backend.close(j.scene)
                j.scene = None
                used -= j.bytes
                j.done = clock() - t0
                j.error = "neighbor overflow: a particle had more neighbors than max_neighbors"
                if log:
                    log(f"[serve] failed req {j.req.rid} {METHOD_NAMES[j.req.method]} n={j.n}: {j.error}")
            js = [j for j in js if j.error is None]
        for j in js:
            j.remaining -= k
        fin = [j for j in js if j.remaining == 0]
        if fin:
            if ln.batch is not None:
                backend.close(ln.batch)
            ln.batch = None
            for j in fin:
                res = read[id(j)]
                j.result = res if keep_results else None
                backend.close(j.scene)
                j.scene = None
                used -= j.bytes
                j.done = clock() - t0
                ln.jobs.remove(j)
                if log:
                    log(f"[serve] done req {j.req.rid} {METHOD_NAMES[j.req.method]} n={j.n} "
                        f"steps={j.req.steps} latency {j.done - j.req.arrival:.3f}s")

    try:
        while pending or waiting or any(ln.jobs for ln in lanes.values()):
            now = clock() - t0
            while pending and pending[0].arrival <= now:
                j = Job(pending.popleft())
                j.spec_future = pool.submit(prepare, j.req)
                waiting.append(j)
                jobs.append(j)
            # with nothing resident there is nothing to overlap: wait for the
            # head request's geometry
            admit(block=not any(ln.jobs for ln in lanes.values()))
            # one quantum at a time per group; of a group's lanes, the one
            # holding the oldest request goes first (large scenes fill the
            # device on their own, so running their work in arrival order
            # finishes each sooner than taking turns would, at the same
            # throughput), then the one that launched longest ago
            busy = {ln.group for ln in lanes.values() if ln.flight is not None}
            order = sorted((ln for ln in lanes.values() if ln.jobs),
                           key=lambda x: (min(j.req.arrival for j in x.jobs), x.turn))
            for ln in order:
                if ln.jobs and ln.flight is None and ln.group not in busy:
                    launches += 1
                    ln.turn = launches
                    launch(ln)
                    if runner is None:
                        land(ln)
                    else:
                        busy.add(ln.group)
            flying = {ln.flight[0]: ln for ln in lanes.values() if ln.flight is not None}
            if not flying:
                if not any(ln.jobs for ln in lanes.values()) and pending and not waiting:
                    sleep(max(0.0, pending[0].arrival - (clock() - t0)))
                continue
            timeout = max(0.0, pending[0].arrival - (clock() - t0)) if pending else None
            # a head request whose geometry is being built wakes the loop too
            head = [waiting[0].spec_future] if waiting and waiting[0].scene is None else []
            done, _ = wait(list(flying) + head, timeout=timeout, return_when=FIRST_COMPLETED)
            for f in done:
                if f in flying:
                    land(flying[f])
    finally:
        if runner is not None:
            # every running quantum ends before its batch is closed; a
            # quantum that completed hands its batch back for closing
            for ln in lanes.values():
                if ln.flight is not None:
                    fut = ln.flight[0]
                    wait([fut])
                    if fut.exception() is None:
                        ln.batch = fut.result()[0]
        for ln in lanes.values():
            if ln.batch is not None:
                backend.close(ln.batch)
            for j in ln.jobs:
                if j.scene is not None:
                    backend.close(j.scene)
        # scenes prepared for requests that were never admitted
        pool.shutdown(wait=True)
        for j in waiting:
            if j.scene is None and j.spec_future.done() and j.spec_future.exception() is None:
                j.scene = j.spec_future.result()[3]
            if j.scene is not None:
                backend.close(j.scene)
        if runner is not None:
            runner.shutdown(wait=True)
    done = [j.done for j in jobs if j.done >= 0]
    first = min(q.arrival for q in requests) if requests else 0.0
    return ServeResult("batched", jobs, (max(done) - first) if done else 0.0, rejected)


def serve_plain(backend, requests: List[Request], budget_bytes: int,
                clock: Callable[[], float] = time.perf_counter, sleep: Callable[[float], None] = time.sleep,
                keep_results: bool = True, log: Callable[[str], None] = None) -> ServeResult:
    """The same requests one at a time, in arrival order, each scene alone."""
    order = sorted(requests, key=lambda q: q.arrival)
    jobs: List[Job] = []
    rejected: List[int] = []
    pool = _spec_pool()
    t0 = clock()
    try:
        futures = {}
        for q in order:
            now = clock() - t0
            if q.arrival > now:
                sleep(q.arrival - now)
            # the geometry of the next arrived requests is built meanwhile
            for q2 in order:
                if q2.arrival <= clock() - t0 and q2.rid not in futures:
                    futures[q2.rid] = pool.submit(build_spec, q2)
            if q.rid not in futures:
                futures[q.rid] = pool.submit(build_spec, q)
            j = Job(q)
            jobs.append(j)
            spec = futures.pop(q.rid).result()
            j.n = spec.n
            scene = backend.create(spec)
            j.bytes = backend.plan_bytes(scene)
            if j.bytes > budget_bytes:
                backend.close(scene)
                rejected.append(q.rid)
                continue
            backend.init(scene)
            j.admitted = clock() - t0
            try:
                backend.step(scene, q.steps)
            except sphg.SphgError as e:
                if not _is_overflow(e):
                    backend.close(scene)
                    raise
                j.error = "neighbor overflow: a particle had more neighbors than max_neighbors"
            if j.error is None:
                res = backend.result(scene)
                j.result = res if keep_results else None
            backend.close(scene)
            j.done = clock() - t0
            if log:
                log(f"[plain] done req {q.rid} {METHOD_NAMES[q.method]} n={j.n} steps={q.steps} "
                    f"latency {j.done - q.arrival:.3f}s")
    finally:
        pool.shutdown(wait=True)
    done = [j.done for j in jobs if j.done >= 0]
    first = min(q.arrival for q in requests) if requests else 0.0
    return ServeResult("plain", jobs, (max(done) - first) if done else 0.0, rejected)


def compare(ctx, reqs: List[Request], budget_bytes: int, quantum: int = 25, check: int = 4,
            parallel: bool = True, split_n: int = SPLIT_N, log: Callable[[str], None] = None,
            large_cap: int = LARGE_CAP) -> dict:
    """Serves `reqs` with continuous batching, then plain (one at a time),
    on the same device; returns both summaries, the throughput ratio and how
    many of the first `check` requests had identical results in both."""
    be = GpuBackend(ctx, lane_contexts=parallel)
    try:
        rb = serve_batched(be, reqs, budget_bytes, quantum, log=log, parallel=parallel, split_n=split_n,
                           large_cap=large_cap)
        rp = serve_plain(be, reqs, budget_bytes, log=log)
    finally:
        be.shutdown()
    sb, sp = rb.summary(), rp.summary()
    same = checked = 0
    byid_b = {j.req.rid: j for j in rb.jobs}
    for j in rp.jobs[:check]:
        jb = byid_b.get(j.req.rid)
        if jb is not None and jb.result is not None and j.result is not None:
            checked += 1
            same += int(np.array_equal(jb.result, j.result))
    ratio = (round(sb["particle_steps_per_s"] / sp["particle_steps_per_s"], 3)
             if sb["particle_steps_per_s"] and sp["particle_steps_per_s"] else None)
    out = {"batched": sb, "plain": sp, "throughput_ratio": ratio, "identical_results": f"{same}/{checked}"}
    bound = throughput_bound(rp)
    if bound:
        out["throughput_bound"] = bound
    return out


def throughput_bound(plain: ServeResult) -> Optional[dict]:
    """The largest batched/plain throughput ratio the stream allows, from
    the plain run: batching cannot step a scene faster than the device's
    saturated particle-step rate, taken per method as the fastest alone rate
    (particle-steps per second from admission to result) of any request of
    that method.  With every request at that rate the stream would take
    sum(work_i / rate[method_i]); the plain makespan over that is the bound.
    A stream whose work sits in scenes already large enough to saturate the
    device alone has a bound near 1."""
    done = [j for j in plain.jobs if j.done >= 0 and j.error is None and j.done > j.admitted]
    if not done or plain.makespan <= 0:
        return None
    best = {}
    for j in done:
        r = j.n * j.req.steps / (j.done - j.admitted)
        best[j.req.method] = max(best.get(j.req.method, 0.0), r)
    ideal = sum(j.n * j.req.steps / best[j.req.method] for j in done)
    big = sum(j.n * j.req.steps for j in done if j.n >= 262144) / sum(j.n * j.req.steps for j in done)
    return {"ratio": round(plain.makespan / ideal, 3),
            "alone_rate_per_method": {METHOD_NAMES[m]: round(r, 1) for m, r in sorted(best.items())},
            "work_share_n_ge_262144": round(big, 3)}


def main():
    ap = argparse.ArgumentParser(description="Serving benchmark: continuous batching vs plain, one GPU")
    ap.add_argument("--requests", type=int, default=48)
    ap.add_argument("--rate", type=float, default=2.0, help="Poisson arrival rate, requests per second")
    ap.add_argument("--nmin", type=int, default=2048)
    ap.add_argument("--nmax", type=int, default=262144)
    ap.add_argument("--smin", type=int, default=100)
    ap.add_argument("--smax", type=int, default=500)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--quantum", type=int, default=25)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--check", type=int, default=4, help="requests whose batched and plain results are compared")
    ap.add_argument("--serial", action="store_true",
                    help="step the lanes' batches one after another on one stream")
    ap.add_argument("--large-cap", type=int, default=LARGE_CAP,
                    help="fluid particles a large-scene lane steps at once, oldest scenes first, at least one "
                         "(0: all it holds)")
    ap.add_argument("--split-n", type=int, default=SPLIT_N,
                    help="fluid size from which a scene steps in its method's large-scene lane (0: one lane "
                         "per method)")
    args = ap.parse_args()

    reqs = request_stream(args.requests, args.rate, args.nmin, args.nmax, args.smin, args.smax, args.seed)
    ctx = sphg.Context(0)
    log = lambda s: print(s, flush=True)   # noqa: E731
    r = compare(ctx, reqs, int(args.budget_gb * 2 ** 30), args.quantum, args.check,
                parallel=not args.serial, split_n=args.split_n, log=log, large_cap=args.large_cap)
    out = {"metric": "served-scene p50/p99 latency and particle-steps/s, continuous batching vs plain GPU",
           "stream": {"requests": args.requests, "rate_per_s": args.rate, "n": [args.nmin, args.nmax],
                      "steps": [args.smin, args.smax], "methods": "WCSPH/PCISPH/DFSPH uniform",
                      "geometry": "dam-break block or sphere drop, spacing 0.02", "seed": args.seed,
                      "budget_gb": args.budget_gb, "quantum": args.quantum,
                      "lanes_in_parallel": not args.serial, "split_n": args.split_n,
                      "large_cap": args.large_cap}}
    out.update(r)
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
