"""Continuous batching on the device: every request's result is the same
alone (plain serving) and in the mix (batched serving), bit for bit."""
import numpy as np
import pytest

from sph_amd import configs, serve, sphg

pytestmark = pytest.mark.gpu


def test_persistent_batch_quanta_match_alone(ctx):
    specs = [configs.small_scene(sphg.DFSPH, n_side=n, seed=s) for n, s in ((6, 11), (9, 12))]
    alone = [sphg.Scene(ctx, s.params, s.fluid, s.boundary) for s in specs]
    for a in alone:
        a.step(9)
    mixed = [sphg.Scene(ctx, s.params, s.fluid, s.boundary) for s in specs]
    b = sphg.Batch(ctx, mixed)
    for k in (2, 3, 4):              # 9 steps in three quanta
        b.run(k)
    b.close()
    for a, m in zip(alone, mixed):
        assert np.array_equal(a.get(sphg.FIELD_X), m.get(sphg.FIELD_X))
        assert np.array_equal(a.get(sphg.FIELD_V), m.get(sphg.FIELD_V))
    for s in alone + mixed:
        s.close()


def test_plan_bytes_matches_allocation(ctx):
    spec = configs.small_scene(sphg.PCISPH, n_side=8)
    s = sphg.Scene(ctx, spec.params, spec.fluid, spec.boundary, init=False)
    planned = s.plan_bytes()
    s.init()
    assert s.info()["device_bytes"] == planned > 0
    s.close()


def test_scenes_prepared_on_another_thread_while_the_context_runs(ctx):
    """Creating, planning and destroying scenes that are not initialised
    touch host memory only and may run on another thread while the context
    steps (the admission of a serving lane does this): the stepped scene
    ends bit-identical to the same scene stepped with nothing beside it."""
    import threading
    spec = configs.small_scene(sphg.DFSPH, n_side=12)
    other = configs.small_scene(sphg.WCSPH, n_side=8)

    def run():
        g = sphg.Scene(ctx, spec.params, spec.fluid, spec.boundary)
        g.step(300)
        ctx.synchronize()
        x, v = g.get(sphg.FIELD_X), g.get(sphg.FIELD_V)
        g.close()
        return x, v

    x0, v0 = run()
    stop = threading.Event()
    made = []
    errors = []

    def prepare():
        try:
            while not stop.is_set():
                s = sphg.Scene(ctx, other.params, other.fluid, other.boundary, init=False)
                made.append(s.plan_bytes())
                s.close()
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append(e)

    t = threading.Thread(target=prepare)
    t.start()
    try:
        x1, v1 = run()
    finally:
        stop.set()
        t.join()
    assert not errors, errors
    # the second thread ran during the steps (each plan takes well under a
    # millisecond against ~60 ms of stepping) and every plan agrees
    assert len(made) >= 1 and len(set(made)) == 1
    assert np.array_equal(x0, x1) and np.array_equal(v0, v1)


def test_served_results_same_alone_and_in_mix(ctx):
    reqs = serve.request_stream(6, rate=500.0, nmin=2048, nmax=6000, smin=12, smax=30, seed=21)
    be = serve.GpuBackend(ctx)
    rb = serve.serve_batched(be, reqs, 8 << 30, quantum=5)
    rp = serve.serve_plain(be, reqs, 8 << 30)
    assert not rb.rejected and not rp.rejected
    byid = {j.req.rid: j for j in rb.jobs}
    methods = set()
    for j in rp.jobs:
        jb = byid[j.req.rid]
        assert j.result is not None and jb.result is not None
        assert np.all(np.isfinite(j.result))
        assert np.array_equal(j.result, jb.result), j.req
        methods.add(j.req.method)
    assert len(methods) >= 2
    # the lanes (method and size class) on their own streams, run concurrently: the same results
    bp = serve.GpuBackend(ctx, lane_contexts=True)
    rc = serve.serve_batched(bp, reqs, 8 << 30, quantum=5, parallel=True, split_n=4096)
    assert not rc.rejected
    byid_c = {j.req.rid: j for j in rc.jobs}
    for j in rp.jobs:
        assert np.array_equal(j.result, byid_c[j.req.rid].result), j.req
    bp.shutdown()


def test_served_stream_with_a_million_particle_request(ctx):
    """A stream shaped like the configured one -- WCSPH/PCISPH/DFSPH, N from
    2k up to a request of ~1.2M particles -- served with lanes by method and
    size on their own streams, and plain: every result is the same alone and
    in the mix, bit for bit, and the large request ran in its own lane."""
    reqs = serve.request_stream(6, rate=400.0, nmin=2048, nmax=20000, smin=10, smax=25, seed=31)
    big = serve.Request(rid=99, arrival=0.002, method=sphg.DFSPH, n_target=1_200_000, steps=12,
                        geometry="block", seed=12345)
    reqs = sorted(reqs + [big], key=lambda q: q.arrival)
    be = serve.GpuBackend(ctx, lane_contexts=True)
    try:
        rb = serve.serve_batched(be, reqs, 64 << 30, quantum=5, parallel=True, split_n=serve.SPLIT_N)
    finally:
        be.shutdown()
    rp = serve.serve_plain(serve.GpuBackend(ctx), reqs, 64 << 30)
    assert not rb.rejected and not rp.rejected
    byid = {j.req.rid: j for j in rb.jobs}
    for j in rp.jobs:
        jb = byid[j.req.rid]
        assert j.result is not None and jb.result is not None, j.req
        assert np.array_equal(j.result, jb.result), j.req
    jb = byid[99]
    assert jb.n >= 1_000_000 and jb.lane == serve.lane_of(sphg.DFSPH, jb.n, serve.SPLIT_N)
    assert jb.lane[1] == serve.LARGE
    assert {j.req.method for j in rp.jobs} == set(serve.METHODS)
