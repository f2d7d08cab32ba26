"""Serving scheduler logic on the CPU: a fake backend with a virtual clock
stands in for the device (the GPU path runs in test_gpu_serve.py)."""
import numpy as np
import pytest

from sph_amd import serve, sphg


class _SyncPool:
    """Builds each request's geometry at submission: with the virtual clock
    a background build would finish at an arbitrary virtual time."""

    def submit(self, fn, *args):
        from concurrent.futures import Future
        f = Future()
        f.set_result(fn(*args))
        return f

    def shutdown(self, wait=True):
        pass


@pytest.fixture(autouse=True)
def _sync_specs(monkeypatch):
    monkeypatch.setattr(serve, "_spec_pool", _SyncPool)


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += max(0.0, dt)


class FakeScene:
    def __init__(self, spec, lane=None):
        self.n = spec.n
        self.method = spec.params.method
        self.lane = lane
        self.steps = 0
        self.ready = False
        self.closed = False
        self.overflow = False       # its neighbor lists overflow on every step


class FakeBatch:
    def __init__(self, be, scenes):
        self.be, self.scenes, self.closed = be, scenes, False
        be.batches.append(sorted({(s.method, s.lane) for s in scenes}))
        be.built.append([(s.method, s.lane) for s in scenes])
        be.built_n.append([(s.lane[1], s.n) for s in scenes])

    def run(self, k):
        with self.be.lock:      # batches of different methods may run on threads
            assert not self.closed
            assert all(s.ready and not s.closed for s in self.scenes)
            self.be.quanta.append((len(self.scenes), k))
            for s in self.scenes:
                s.steps += k
            # one launch sequence for the whole batch: fixed cost + work
            self.be.clock.t += k * (2e-4 + 1e-9 * sum(s.n for s in self.scenes))
        if any(s.overflow for s in self.scenes):
            # as the library: every scene stepped, the call reports the overflow
            raise sphg.SphgError(sphg.EOVERFLOW, "neighbor overflow")

    def close(self):
        self.closed = True


class FakeBackend:
    BYTES_PER_PARTICLE = 600

    def __init__(self, clock):
        import threading
        self.lock = threading.Lock()
        self.clock = clock
        self.quanta = []
        self.batches = []       # the (method, lane) pairs of each batch built
        self.built = []         # the same per scene of each batch built
        self.built_n = []       # (size class, fluid particles) per scene of each batch built
        self.used = 0
        self.peak = 0
        self.admitted = []

    def create(self, spec, lane=None):
        return FakeScene(spec, lane)

    def plan_bytes(self, s):
        return s.n * self.BYTES_PER_PARTICLE

    def init(self, s):
        s.ready = True
        self.used += s.n * self.BYTES_PER_PARTICLE
        self.peak = max(self.peak, self.used)
        self.admitted.append(s)

    def batch(self, scenes):
        return FakeBatch(self, scenes)

    def step(self, s, k):
        s.steps += k
        self.clock.t += k * (2e-4 + 1e-9 * s.n)

    def result(self, s):
        return np.array([s.steps], dtype=np.int64)

    def overflowed(self, s):
        return s.overflow

    def close(self, obj):
        if isinstance(obj, FakeScene) and obj.ready and not obj.closed:
            self.used -= obj.n * self.BYTES_PER_PARTICLE
        obj.closed = True


def _stream(count=24, rate=50.0, seed=9):
    return serve.request_stream(count, rate, nmin=512, nmax=4096, smin=10, smax=60, seed=seed)


def test_request_stream_is_seeded_and_in_range():
    a = serve.request_stream(200, 3.0, seed=5)
    b = serve.request_stream(200, 3.0, seed=5)
    assert [(q.arrival, q.method, q.n_target, q.steps, q.geometry) for q in a] == \
           [(q.arrival, q.method, q.n_target, q.steps, q.geometry) for q in b]
    assert all(2048 <= q.n_target <= 2_000_000 and 100 <= q.steps <= 2000 for q in a)
    assert all(x.arrival < y.arrival for x, y in zip(a, a[1:]))
    assert {q.method for q in a} == set(serve.METHODS)
    assert {q.geometry for q in a} == {"block", "drop"}
    # mean inter-arrival ~ 1/rate
    gaps = np.diff([0.0] + [q.arrival for q in a])
    assert 0.25 < gaps.mean() < 0.42


def test_build_spec_sizes():
    for q in _stream(12):
        spec = serve.build_spec(q)
        assert spec.n > 0 and spec.nb > 0
        assert 0.5 * q.n_target <= spec.n <= 1.6 * q.n_target, (q, spec.n)
        assert spec.params.method == q.method


def test_batched_exact_steps_budget_fifo_quantum():
    clk = Clock()
    be = FakeBackend(clk)
    reqs = _stream()
    budget = 3 * 4096 * FakeBackend.BYTES_PER_PARTICLE   # only a few scenes fit at once
    res = serve.serve_batched(be, reqs, budget, quantum=7, clock=clk, sleep=clk.sleep)
    assert not res.rejected
    assert len(res.jobs) == len(reqs)
    for j in res.jobs:
        assert j.done >= 0 and j.result is not None
        assert int(j.result[0]) == j.req.steps          # exactly the requested steps
        assert j.admitted >= j.req.arrival - 1e-12      # never before it arrived
        assert j.done >= j.admitted
    assert be.peak <= budget
    assert be.used == 0                                  # everything freed
    assert all(k <= 7 for _, k in be.quanta)
    # first come, first served: admission order is arrival order
    order = [j.req.rid for j in sorted(res.jobs, key=lambda j: j.admitted)]
    assert order == sorted(order)
    # scenes actually shared launches
    assert max(n for n, _ in be.quanta) > 1


def test_batched_lanes_in_parallel_same_steps_and_budget():
    """The lanes' batches run their quanta on threads, each lane refilled as
    soon as its own quantum ends: every request still gets exactly its
    steps, admission stays within the budget and in arrival order, and
    everything is freed."""
    for split in (0, 2048):
        clk = Clock()
        be = FakeBackend(clk)
        reqs = _stream()
        budget = 3 * 4096 * FakeBackend.BYTES_PER_PARTICLE
        res = serve.serve_batched(be, reqs, budget, quantum=7, clock=clk, sleep=clk.sleep, parallel=True,
                                  split_n=split)
        assert not res.rejected and len(res.jobs) == len(reqs)
        for j in res.jobs:
            assert int(j.result[0]) == j.req.steps
            assert j.admitted >= j.req.arrival - 1e-12
        assert be.peak <= budget and be.used == 0
        order = [j.req.rid for j in sorted(res.jobs, key=lambda j: j.admitted)]
        assert order == sorted(order)


def test_lanes_split_by_method_and_size():
    """With a size boundary every batch holds scenes of one method and one
    size class, and both classes occur; without one, one lane per method."""
    reqs = _stream(30, rate=2000.0)
    for split, parallel in ((2048, False), (2048, True), (0, False)):
        clk = Clock()
        be = FakeBackend(clk)
        res = serve.serve_batched(be, reqs, 1 << 40, quantum=5, clock=clk, sleep=clk.sleep, parallel=parallel,
                                  split_n=split)
        assert all(int(j.result[0]) == j.req.steps for j in res.jobs)
        assert all(len(b) == 1 for b in be.batches)
        for j in res.jobs:
            assert j.lane == serve.lane_of(j.req.method, j.n, split)
        classes = {b[0][1][1] for b in be.batches}
        assert classes == ({serve.SMALL, serve.LARGE} if split else {serve.SMALL})


def test_batched_rejects_what_never_fits():
    clk = Clock()
    be = FakeBackend(clk)
    reqs = _stream(6)
    budget = 100 * FakeBackend.BYTES_PER_PARTICLE        # smaller than any scene
    res = serve.serve_batched(be, reqs, budget, clock=clk, sleep=clk.sleep)
    assert sorted(res.rejected) == [q.rid for q in reqs]
    assert be.used == 0


def test_plain_is_sequential_and_exact():
    clk = Clock()
    be = FakeBackend(clk)
    reqs = _stream(10)
    res = serve.serve_plain(be, reqs, 1 << 40, clock=clk, sleep=clk.sleep)
    done = [j for j in res.jobs if j.done >= 0]
    assert len(done) == len(reqs)
    for a, b in zip(done, done[1:]):
        assert b.admitted >= a.done                     # one scene at a time
    for j in done:
        assert int(j.result[0]) == j.req.steps
    s = res.summary()
    assert s["requests"] == len(reqs) and s["particle_steps_per_s"] > 0


def test_batching_beats_plain_on_launch_bound_stream():
    """With a per-launch cost that dominates small scenes, sharing launches
    raises throughput and lowers latency (the fake cost model)."""
    reqs = _stream(30, rate=2000.0)   # arrivals faster than one scene runs
    c1, c2 = Clock(), Clock()
    rb = serve.serve_batched(FakeBackend(c1), reqs, 1 << 40, quantum=25, clock=c1, sleep=c1.sleep)
    rp = serve.serve_plain(FakeBackend(c2), reqs, 1 << 40, clock=c2, sleep=c2.sleep)
    sb, sp = rb.summary(), rp.summary()
    assert sb["particle_steps_per_s"] > 2 * sp["particle_steps_per_s"]
    assert sb["latency_p50_s"] < sp["latency_p50_s"]


def test_neighbor_overflow_fails_only_its_request():
    """A scene whose neighbor lists overflow fails its own request; the other
    resident scenes of its batch and of the other methods finish with their
    exact step counts, and the memory comes back."""
    for parallel in (False, True):
        clk = Clock()
        be = FakeBackend(clk)
        reqs = _stream(16, rate=200.0, seed=11)
        bad = reqs[3].rid
        create = be.create

        def create_marked(spec, lane=None, _c=create):
            sc = _c(spec, lane)
            sc.overflow = spec.name.startswith(f"req{bad}_")
            return sc
        be.create = create_marked
        r = serve.serve_batched(be, reqs, budget_bytes=10 ** 12, quantum=7, clock=clk, sleep=clk.sleep,
                                parallel=parallel, split_n=2048 if parallel else 0)
        byid = {j.req.rid: j for j in r.jobs}
        assert byid[bad].error and byid[bad].result is None
        for q in reqs:
            if q.rid != bad:
                j = byid[q.rid]
                assert j.error is None and int(j.result[0]) == q.steps, (q.rid, j.error)
        s = r.summary()
        assert s["failed"] == 1 and s["requests"] == len(reqs) - 1
        assert be.used == 0

    clk = Clock()
    be = FakeBackend(clk)
    reqs = _stream(6, rate=200.0, seed=12)
    bad = reqs[2].rid
    create = be.create

    def create_marked2(spec, lane=None, _c=create):
        sc = _c(spec, lane)
        sc.overflow = spec.name.startswith(f"req{bad}_")
        return sc
    be.create = create_marked2

    def step(sc, k, _s=be.step):
        _s(sc, k)
        if sc.overflow:
            raise sphg.SphgError(sphg.EOVERFLOW, "neighbor overflow")
    be.step = step
    r = serve.serve_plain(be, reqs, budget_bytes=10 ** 12, clock=clk, sleep=clk.sleep)
    byid = {j.req.rid: j for j in r.jobs}
    assert byid[bad].error and all(byid[q.rid].error is None for q in reqs if q.rid != bad)
    assert r.summary()["failed"] == 1


def test_throughput_bound_from_plain_run():
    """The bound divides the plain makespan by the stream's time at each
    method's fastest alone rate: with a cost model whose per-particle rate
    rises with n, the bound lies above 1 and below the rate spread."""
    clk = Clock()
    be = FakeBackend(clk)
    reqs = _stream(12, rate=2000.0)
    rp = serve.serve_plain(be, reqs, 1 << 40, clock=clk, sleep=clk.sleep)
    b = serve.throughput_bound(rp)
    assert b is not None and set(b["alone_rate_per_method"]) <= {"WCSPH", "PCISPH", "DFSPH"}
    # fake step cost 2e-4 + 1e-9 n per step: the rate n / (2e-4 + 1e-9 n) grows with n
    ns = [j.n for j in rp.jobs]
    spread = (max(ns) / (2e-4 + 1e-9 * max(ns))) / (min(ns) / (2e-4 + 1e-9 * min(ns)))
    assert 1.0 < b["ratio"] <= spread * 1.001
    assert b["work_share_n_ge_262144"] == 0.0


def test_admission_does_not_wait_for_geometry_while_scenes_run(monkeypatch):
    """A request whose geometry is still being built (here: ready 0.05 s of
    virtual time after its submission) does not hold up the resident
    scenes: their quanta go on, and the request is admitted at the first
    turn of the loop after its geometry is ready.  With nothing resident
    the loop waits for it."""
    from concurrent.futures import Future
    clk = Clock()
    delay = 0.05

    class LateFuture(Future):
        def result(self, timeout=None):
            if not self.done():          # waiting for the build: the clock moves on to it
                clk.t = max(clk.t, self.ready)
                self.set_result(self.fn(*self.args))
            return super().result()

    class DelayedPool:
        pending = []

        def submit(self, fn, *args):
            f = LateFuture()
            f.fn, f.args, f.ready = fn, args, clk.t + delay
            DelayedPool.pending.append(f)
            return f

        def shutdown(self, wait=True):
            pass

    def tick():
        for f in DelayedPool.pending:
            if not f.done() and clk.t >= f.ready:
                f.set_result(f.fn(*f.args))

    monkeypatch.setattr(serve, "_spec_pool", DelayedPool)
    be = FakeBackend(clk)
    run = FakeBatch.run

    starts = []

    def run_and_tick(self, k):
        starts.append(clk.t)
        run(self, k)
        tick()
    monkeypatch.setattr(FakeBatch, "run", run_and_tick)
    reqs = serve.request_stream(6, 200.0, nmin=512, nmax=4096, smin=200, smax=300, seed=3)
    res = serve.serve_batched(be, reqs, 1 << 40, quantum=5, clock=clk, sleep=clk.sleep)
    byid = {j.req.rid: j for j in res.jobs}
    ready = {f.args[0].rid: f.ready for f in DelayedPool.pending}
    for q in reqs:
        j = byid[q.rid]
        assert int(j.result[0]) == q.steps
        assert j.admitted >= ready[q.rid] - 1e-12              # not before its geometry is ready
        assert ready[q.rid] >= q.arrival + delay - 1e-12       # built from its arrival on
    # the first request ran alone while the others' geometry was built, and
    # each later one joined within a quantum of its geometry being ready
    assert be.quanta[0][0] == 1
    for q in reqs[1:]:
        assert byid[q.rid].admitted - ready[q.rid] < 0.01, q.rid
    # quanta ran while the second request's geometry was being built (a
    # blocking admission would have stopped the clock's other work there)
    r1 = ready[reqs[1].rid]
    assert any(r1 - delay < t < r1 for t in starts)


def test_scene_initialised_only_while_its_lane_is_idle():
    """A scene is initialised on its lane group's context, and batches are
    built and closed there, never while a quantum of the same group runs (a
    context is used by one host thread at a time); a group runs one quantum
    at a time, the large lanes of all methods share a group and every small
    lane has its own.  The quanta run on threads and take a little real time
    here, so an overlap would show."""
    import threading
    import time as _time
    clk = Clock()
    be = FakeBackend(clk)
    running = set()
    guard = threading.Lock()
    run = FakeBatch.run
    lanes_seen = []
    groups_seen = set()

    def run_slowly(self, k):
        lanes = {s.lane for s in self.scenes}
        with guard:
            # one quantum at a time per context
            assert not lanes & running, (lanes, running)
            running.update(lanes)
            groups_seen.update((s.method, s.lane) for s in self.scenes)
        try:
            _time.sleep(0.002)
            run(self, k)
        finally:
            with guard:
                running.difference_update(lanes)

    init, batch, close = be.init, be.batch, be.close

    def init_checked(s):
        with guard:
            assert s.lane not in running, s.lane
            lanes_seen.append(s.lane)
        init(s)

    def batch_checked(scenes):
        with guard:
            assert not {s.lane for s in scenes} & running
        return batch(scenes)

    def close_checked(obj):
        # an initialised scene or a batch is closed only while its lane is idle
        lanes = {s.lane for s in obj.scenes} if isinstance(obj, FakeBatch) else (
            {obj.lane} if obj.ready else set())
        with guard:
            assert not lanes & running, lanes
        close(obj)
    be.init, be.batch, be.close = init_checked, batch_checked, close_checked
    FakeBatch.run = run_slowly
    try:
        reqs = _stream(20, rate=300.0, seed=13)
        res = serve.serve_batched(be, reqs, 1 << 40, quantum=3, clock=clk, sleep=clk.sleep, parallel=True,
                                  split_n=2048)
    finally:
        FakeBatch.run = run
    assert all(int(j.result[0]) == j.req.steps for j in res.jobs)
    assert len(lanes_seen) == len(reqs) and len(set(lanes_seen)) > 2
    assert be.used == 0
    # the scenes' contexts: the large lanes of all methods share one, each
    # small lane has its own (at most 4 for 3 methods)
    large = {m for m, g in groups_seen if g == (None, serve.LARGE)}
    assert len(large) > 1
    assert len({g for _, g in groups_seen}) <= 4
    for j in res.jobs:
        assert j.group == serve.group_of(j.lane, 2048)


def test_parallel_lanes_need_a_context_per_lane():
    """Lanes that run on threads of their own need a context each: a backend
    that drives every lane through one context is refused."""
    clk = Clock()
    be = FakeBackend(clk)
    be.lane_contexts = False
    with pytest.raises(ValueError, match="context per lane"):
        serve.serve_batched(be, _stream(4), 1 << 40, clock=clk, sleep=clk.sleep, parallel=True)
    res = serve.serve_batched(be, _stream(4), 1 << 40, clock=clk, sleep=clk.sleep, parallel=False)
    assert all(int(j.result[0]) == j.req.steps for j in res.jobs)


def test_large_lanes_run_oldest_request_first():
    """The large lanes of all methods share a context; of them, the lane
    holding the oldest request runs until that request is done, so large
    requests complete in arrival order, whatever their step counts (taking
    turns would finish the shortest first)."""
    clk = Clock()
    be = FakeBackend(clk)
    steps = (300, 200, 100)
    reqs = [serve.Request(i, 0.001 * i, m, 4096, s, "block", 10_000 + i)
            for i, (m, s) in enumerate(zip((sphg.DFSPH, sphg.PCISPH, sphg.WCSPH), steps))]
    res = serve.serve_batched(be, reqs, 1 << 40, quantum=10, clock=clk, sleep=clk.sleep, parallel=True,
                              split_n=2048)
    assert all(int(j.result[0]) == j.req.steps for j in res.jobs)
    assert all(j.group == (None, serve.LARGE) for j in res.jobs)
    done = [j.req.rid for j in sorted(res.jobs, key=lambda j: j.done)]
    assert done == [0, 1, 2]


def test_large_lane_steps_its_oldest_scenes_only():
    """With large_cap a large lane's batches hold its oldest scenes up to
    that many fluid particles (at least one scene); the others wait resident
    (their memory is reserved) and every request still gets exactly its
    steps; small lanes are not capped."""
    for cap in (1, 6000):
        clk = Clock()
        be = FakeBackend(clk)
        reqs = _stream(30, rate=2000.0)
        res = serve.serve_batched(be, reqs, 1 << 40, quantum=5, clock=clk, sleep=clk.sleep, parallel=True,
                                  split_n=2048, large_cap=cap)
        assert all(int(j.result[0]) == j.req.steps for j in res.jobs)
        large = [b for b in be.built_n if b[0][0] == serve.LARGE]
        small = [b for b in be.built_n if b[0][0] == serve.SMALL]
        assert large and small
        for b in large:
            ns = [n for _, n in b]
            assert len(ns) == 1 or sum(ns) <= cap
        if cap == 1:
            assert max(len(b) for b in large) == 1
        else:
            assert max(len(b) for b in large) > 1
        assert max(len(b) for b in small) > 1
        assert be.used == 0


def test_failed_quantum_aborts_and_frees_everything():
    """A quantum that fails with an error other than a neighbor overflow
    ends serving with that error; every batch is closed before its scenes,
    and every scene is closed."""
    for parallel in (False, True):
        clk = Clock()
        be = FakeBackend(clk)
        built = []
        batch = be.batch

        def batch_recorded(scenes):
            b = batch(scenes)
            built.append(b)
            return b
        be.batch = batch_recorded
        run = FakeBatch.run
        calls = [0]

        def failing(self, k):
            calls[0] += 1
            if calls[0] == 6:
                raise sphg.SphgError(sphg.EHIP, "device error")
            run(self, k)
        FakeBatch.run = failing
        try:
            with pytest.raises(sphg.SphgError, match="device error"):
                serve.serve_batched(be, _stream(20, rate=500.0), 1 << 40, quantum=5, clock=clk, sleep=clk.sleep,
                                    parallel=parallel, split_n=2048, large_cap=3000)
        finally:
            FakeBatch.run = run
        assert built and all(b.closed for b in built)
        assert be.used == 0
