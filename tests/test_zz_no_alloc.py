"""sphg_batch_step_noalloc against sphg_batch_step: the same scenes stepped
by both entry points end bit-identical, on two calls in a row, the second
with a smaller batch (fewer and smaller scenes, another method) that reuses
the static table array the first call wrote."""
import ctypes as C

import numpy as np
import pytest

from sph_amd import configs, sphg

pytestmark = pytest.mark.gpu


def _scenes(ctx, specs):
    return [sphg.Scene(ctx, s.params, s.fluid, s.boundary) for s in specs]


def _state(s):
    sc = s.scalars()
    return s.get(sphg.FIELD_X), s.get(sphg.FIELD_V), (sc.time, sc.time_step, sc.sum_iterations, sc.sum_iterations_v)


@pytest.mark.parametrize("calls", [[(sphg.DFSPH, (10, 12, 14), 6), (sphg.WCSPH, (8, 9), 4)],
                                   [(sphg.PCISPH, (11, 9, 10), 5), (sphg.DFSPH, (8,), 3)]])
def test_noalloc_batch_step_matches_batch_step(ctx, calls):
    for method, sides, steps in calls:
        specs = [configs.small_scene(method, n_side=n, seed=20 + n) for n in sides]
        a, b = _scenes(ctx, specs), _scenes(ctx, specs)
        print(f"[no-alloc] {method=} sides={sides} steps={steps}", flush=True)
        sa = ctx.batch_step(a, steps)
        sb = ctx.batch_step_noalloc(b, steps)
        assert sa.steps == sb.steps == steps and sa.iterations == sb.iterations
        for ga, gb in zip(a, b):
            xa, va, ka = _state(ga)
            xb, vb, kb = _state(gb)
            assert np.array_equal(xa, xb) and np.array_equal(va, vb) and ka == kb
        for s in a + b:
            s.close()


def test_noalloc_refuses_kernel_timing(ctx):
    spec = configs.small_scene(sphg.WCSPH, n_side=8)
    s = sphg.Scene(ctx, spec.params, spec.fluid, spec.boundary)
    ctx.kernel_timing(True)
    try:
        with pytest.raises(sphg.SphgError) as e:
            ctx.batch_step_noalloc([s], 1)
        assert e.value.code == sphg.ESTATE
    finally:
        ctx.kernel_timing(False)
    ctx.batch_step_noalloc([s], 1)
    s.close()


def test_noalloc_symbol_exported():
    lib = sphg.load()
    assert isinstance(lib.sphg_batch_step_noalloc, C._CFuncPtr)
