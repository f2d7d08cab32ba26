"""Serving scheduler logic on the CPU: a fake backend with a virtual clock
stands in for the device (the GPU path runs in test_gpu_serve.py)."""
import numpy as np

from sph_amd import serve, sphg


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += max(0.0, dt)


class FakeScene:
    def __init__(self, spec):
        self.n = spec.n
        self.steps = 0
        self.ready = False
        self.closed = False
        self.overflow = False       # its neighbor lists overflow on every step


class FakeBatch:
    def __init__(self, be, scenes):
        self.be, self.scenes, self.closed = be, scenes, False

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
        self.used = 0
        self.peak = 0
        self.admitted = []

    def create(self, spec):
        return FakeScene(spec)

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


def test_batched_methods_in_parallel_same_steps_and_budget():
    """The methods' batches run their quanta on threads: every request still
    gets exactly its steps, admission stays within the budget and in arrival
    order, and everything is freed."""
    clk = Clock()
    be = FakeBackend(clk)
    reqs = _stream()
    budget = 3 * 4096 * FakeBackend.BYTES_PER_PARTICLE
    res = serve.serve_batched(be, reqs, budget, quantum=7, clock=clk, sleep=clk.sleep, parallel_methods=True)
    assert not res.rejected and len(res.jobs) == len(reqs)
    for j in res.jobs:
        assert int(j.result[0]) == j.req.steps
        assert j.admitted >= j.req.arrival - 1e-12
    assert be.peak <= budget and be.used == 0
    order = [j.req.rid for j in sorted(res.jobs, key=lambda j: j.admitted)]
    assert order == sorted(order)


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

        def create_marked(spec, _c=create):
            sc = _c(spec)
            sc.overflow = spec.name.startswith(f"req{bad}_")
            return sc
        be.create = create_marked
        r = serve.serve_batched(be, reqs, budget_bytes=10 ** 12, quantum=7, clock=clk, sleep=clk.sleep,
                                parallel_methods=parallel)
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

    def create_marked2(spec, _c=create):
        sc = _c(spec)
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
