"""The benchmark workloads at full size, HIP solver against the FP64 oracle.

Each test runs a configuration at its real size on the device, hands the
device's state to the oracle (oracle_set_state: the positions, velocities,
warm-start accumulators, time step and last iteration count, from which it
rebuilds the grid, densities and alpha), and steps both once from there:

  * DFSPH dam break, 1,000,000 fluid + 102,410 boundary particles, after 150
    steps (the collapse under way, ~15 density iterations per step): cell
    keys of every particle, the buckets (membership and ascending-id order)
    of every cell, and the neighbor sets of 10,000 sampled particles plus
    the 512 with the most neighbors, bit-exact; then one step's fields;
  * PCISPH spheres with Akinci surface tension and adhesion (310,571 fluid,
    247,004 boundary particles) after 20 steps, then one step;
  * WCSPH dam break, 32,768 fluid particles with XSPH, 8 steps, each from
    the device's state;
  * a dense lattice (half the rest spacing, ~260 neighbors per particle),
    whose wavefronts' candidate rows exceed the list build's LDS stage, so
    the build reads part of them from global memory;
  * the 256-scene batched DFSPH workload (17.4M fluid particles), 5 batched
    steps: sampled scenes bit-identical to the same scenes stepped alone,
    and two scenes against the oracle at every step.

Fields are held to 1e-5 relative, normwise (max|gpu - ref| / max|ref|),
the divergence error to 5e-5 against the FP64 run and to 1e-5 against the
oracle's divergence error at the device's FP32 state (see TOL_DIVERR);
density deviations rho - rho0 and rho* - rho0 are compared as stored,
against max(max|ref|, eta rho0), and the divergence error against
max(max|ref|, eta_V rho0 / dt).  Iteration counts agree within
one; fields are compared on a step whose counts agree (up to three
consecutive steps are tried).  The DFSPH oracle rounds its updated positions
to FP32 as the device stores them; the error against the FP64 run of the
same step is logged beside.

The reference's DFSPH eps branches (|k_i + k_j| > 1e-5) are hard
thresholds: a particle whose rho* - rho0 puts a pair sum within the FP32
state's resolution of 1e-5 may take the other branch on the device, which
moves it and its neighbors by ~1e-4 m/s.  Such particles are found by
stepping the oracle with the threshold moved by -1e-4 and +1e-4 (relative);
they are counted (~2,100 of the 1M particles in one step, with the
neighbors their flips move; at most 0.5% is allowed) and left out of the
field comparison.
"""
import os
import time

import numpy as np
import pytest

from conftest import dev_err, rel_norm
from sph_amd import configs, sphg
import oracle.pyoracle as po

pytestmark = pytest.mark.gpu

TOL = 1e-5
# The divergence error read back after the divergence solve is a cancelling
# sum over velocities that the device stores in FP32 after every divergence
# sweep (the oracle keeps them in FP64): at 1M particles, speeds up to 3 m/s,
# the rounding (6e-8 relative) leaves ~2e-5 of its largest value (measured
# 2.3e-5).  It is therefore checked twice: against the FP64 run at 5e-5, and
# at 1e-5 against the oracle's divergence error at the device's own FP32
# positions and velocities (`diverr_at_state`), which isolates the function
# from the rounding of its inputs; the velocities themselves and kappa_v,
# which the solver takes from the divergence error, are within 1e-5 of the
# FP64 run.
TOL_DIVERR = 5e-5


def _log(msg):
    print(f"[configs] {msg}", flush=True)


def _state(g):
    sc = g.scalars()
    dfs = g.params.method == sphg.DFSPH
    return (g.get(sphg.FIELD_X), g.get(sphg.FIELD_V), g.get(sphg.FIELD_K) if dfs else None,
            g.get(sphg.FIELD_KV) if dfs else None, sc.time_step, sc.iterations)


def _inject(g, o, state=None):
    """Hands the device's state (or a saved one) to the oracle.  A DFSPH
    oracle rounds its updated positions to FP32 (oracle_set_f32_positions),
    as the device stores them, so the grid rebuild, densities, alpha and the
    divergence solve of the step see the same positions on both sides."""
    o.set_state(*(state or _state(g)))
    if g.params.method == sphg.DFSPH:
        o.set_f32_positions(True)


def _floor(p):
    """Deviation scale below which density errors are measured against the
    solver's tolerance eta rho0 (a scene near rest has deviations near 0)."""
    return p.eta * p.density0


_DF = {"v": (sphg.FIELD_V, po.OF_V), "x": (sphg.FIELD_X, po.OF_X), "drho": (sphg.FIELD_DRHO, po.OF_RHO),
       "drhostar": (sphg.FIELD_DRHOSTAR, po.OF_RHOSTAR), "alpha": (sphg.FIELD_ALPHA, po.OF_ALPHA),
       "k": (sphg.FIELD_K, po.OF_K), "kv": (sphg.FIELD_KV, po.OF_KV), "diverr": (sphg.FIELD_DIVERR, po.OF_DIVERR)}


def _oracle_fields(o, rho0):
    out = {k: o.get(f) for k, (_, f) in _DF.items()}
    out["drho"] = out["drho"] - rho0
    out["drhostar"] = out["drhostar"] - rho0
    return out


def _per(a):
    a = np.abs(np.asarray(a, dtype=np.float64))
    return a.max(1) if a.ndim > 1 else a


def _eps_sensitive(o, start, ref, rho0, rel=1e-4):
    """Particles whose step depends on a DFSPH eps branch (|k_i + k_j| >
    1e-5, src/DFSPHSolver.cpp:26,364,642) decided within `rel` of its
    threshold: the oracle steps again from `start` with the threshold moved
    by -rel and by +rel, and a particle is sensitive where either run moves
    one of its fields by more than 1e-7 of the field's largest value.  An
    FP32 state resolves rho* - rho0 to ~1e-9 kg/m^3, so near the threshold
    its decision may go the other way than the FP64 run's; such a flip moves
    the velocities of the particle and its neighbors by ~1e-4 m/s."""
    sens = None
    for f in (1.0 - rel, 1.0 + rel):
        o.set_state(*start)
        o.set_f32_positions(True)
        o.set_eps(1e-5 * f)
        o.step(1)
        alt = _oracle_fields(o, rho0)
        for k, r in ref.items():
            m = _per(alt[k] - r) > 1e-7 * max(float(np.abs(r).max()), 1e-30)
            sens = m if sens is None else (sens | m)
    o.set_eps(1e-5)
    return sens


def _dfsph_fields(g, o, mask=None, ref=None):
    """Normwise errors of the DFSPH fields; particles in `mask` are left out
    (normalised by the whole field's largest value all the same).  Density
    deviations are normalised by at least eta rho0 and the divergence error
    by at least eta_V rho0 / dt, the solves' own tolerances."""
    rho0, fl = g.params.density0, _floor(g.params)
    fl_div = g.params.eta_v * rho0 / o.scalars()["time_step"]
    ref = ref or _oracle_fields(o, rho0)
    keep = slice(None) if mask is None else ~mask
    out = {}
    for k, (gf, _) in _DF.items():
        a = np.asarray(g.get(gf), dtype=np.float64)
        b = ref[k]
        den = float(np.abs(b).max())
        if k in ("drho", "drhostar"):
            den = max(den, fl)
        if k == "diverr":
            den = max(den, fl_div)
        num = float(_per(a - b)[keep].max()) if a.size else 0.0
        out[k] = num / den if den > 0 else num
    return out


def _diverr_at_state(g, o):
    """The device's divergence error against the oracle's evaluation of the
    same function at the device's own FP32 positions and velocities (the
    oracle's grid, lists and densities are rebuilt from those positions; its
    state is replaced), over every particle: the function itself, apart from
    the rounding of its inputs that the oracle's run does not share.
    Normalised by max(max|ref|, eta_V rho0 / dt) like the FP64 comparison;
    the plain normwise value is returned beside it."""
    dt = o.scalars()["time_step"]
    o.set_state(g.get(sphg.FIELD_X), g.get(sphg.FIELD_V), None, None, dt, 0)
    ref = o.divergence_error(g.get(sphg.FIELD_V))
    got = g.get(sphg.FIELD_DIVERR)
    fl = g.params.eta_v * g.params.density0 / dt
    return dev_err(got, ref, 0.0, fl), rel_norm(got, ref)


def _check(errs, tols=None):
    tols = tols or {}
    bad = {k: v for k, v in errs.items() if not v <= tols.get(k, TOL)}
    assert not bad, bad


def _grid_parity(g, o, spec, sample=10000, dense=512, seed=5):
    """Keys, buckets and sampled neighbor sets of the device's current grid
    against the oracle's grid of the same positions."""
    okeys = o.keys()
    assert np.array_equal(g.get(sphg.FIELD_KEYS), okeys), "cell keys"
    ids, cells = g.sorted_ids()
    assert np.array_equal(np.sort(ids), np.arange(spec.n)), "sorted ids are a permutation"
    assert np.all(np.diff(cells) >= 0), "cells sorted"
    same = np.diff(cells) == 0
    assert np.all(np.diff(ids.astype(np.int64))[same] > 0), "within-cell order is ascending id"
    # every run of equal device cell is exactly one reference bucket
    ks = okeys[ids]
    assert np.array_equal(~same, (np.diff(ks, axis=0) != 0).any(1)), "bucket membership"
    counts, nb = g.neighbors()
    ngrad = g.get(sphg.FIELD_NGRAD)
    rng = np.random.default_rng(seed)
    pick = np.unique(np.concatenate([rng.choice(spec.n, size=min(sample, spec.n), replace=False),
                                     np.argsort(counts)[-dense:]]))
    x64 = g.get(sphg.FIELD_X).astype(np.float64)
    h = spec.params.support_radius
    for i in pick:
        cand = o.candidates(int(i))
        d = x64[cand] - x64[i]
        rl = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
        q = rl / h
        keep = (q <= 1.0) & (cand != i)
        got = nb[i, : counts[i]]
        assert set(got.tolist()) == set(cand[keep].tolist()), f"neighbors of particle {i}"
        ql, rk = q[keep], rl[keep]
        nz = (rk > 1e-5) & np.where(ql <= 0.5, ql * (3.0 * ql - 2.0) != 0.0, (1.0 - ql) * (1.0 - ql) != 0.0)
        # the stored count saturates at 255 (it feeds the cutoff at 20)
        assert ngrad[i] == min(int(nz.sum()), 255), f"nonzero-gradient count of particle {i}"
    return len(pick), int(counts.max())


def test_dfsph_dam_1m_from_mid_collapse(ctx):
    t0 = time.time()
    spec = configs.dfsph_dam_1m()
    g = sphg.Scene(ctx, spec.params, spec.fluid, spec.boundary)
    g.step(150)
    _log(f"1M: 150 device steps, t = {g.scalars().time:.4f} s, it_d {g.scalars().iterations} ({time.time() - t0:.1f} s)")
    o = po.Oracle(spec.params, spec.fluid, spec.boundary)
    start = _state(g)
    _inject(g, o, start)
    npick, cmax = _grid_parity(g, o, spec)
    _log(f"1M: keys, buckets and {npick} neighbor sets exact (longest list {cmax}) ({time.time() - t0:.1f} s)")
    matched = None
    for attempt in range(3):
        if attempt:
            start = _state(g)
            _inject(g, o, start)
        g.step(1)
        o.step(1)
        sg, so = g.scalars(), o.scalars()
        its = (sg.iterations, int(so["iterations"]), sg.iterations_v, int(so["iterations_v"]))
        assert abs(its[0] - its[1]) <= 1 and abs(its[2] - its[3]) <= 1, its
        assert abs(sg.time_step - so["time_step"]) <= 1e-6 * so["time_step"]
        if its[0] == its[1] and its[2] == its[3]:
            ref = _oracle_fields(o, spec.params.density0)
            sens = _eps_sensitive(o, start, ref, spec.params.density0)
            matched = _dfsph_fields(g, o, sens, ref)
            whole = _dfsph_fields(g, o, None, ref)
            matched["diverr_at_state"], plain = _diverr_at_state(g, o)
            _log(f"1M: step matched, iterations {its}; {int(sens.sum())} particles at an eps branch threshold; "
                 f"errors elsewhere { {k: f'{v:.1e}' for k, v in matched.items()} }, over all particles "
                 f"{ {k: f'{v:.1e}' for k, v in whole.items()} }; divergence error at the device's state, "
                 f"plain normwise {plain:.1e} ({time.time() - t0:.1f} s)")
            break
        _log(f"1M: iteration counts {its} differ by one; next step")
    assert matched is not None
    assert sens.sum() <= 5e-3 * spec.n, int(sens.sum())
    _check(matched, {"diverr": TOL_DIVERR})
    # the same step in FP64 throughout (positions not rounded): the error of
    # the FP32 state against an FP64 run, stated; particles whose count of
    # nonzero-gradient neighbors sits at the divergence cutoff of 20 can
    # switch sides on a one-ulp position difference, so this is not held to
    # the parity bar
    o.set_state(*start)
    o.set_f32_positions(False)
    o.step(1)
    so = o.scalars()
    if (int(so["iterations"]), int(so["iterations_v"])) == (its[0], its[2]):
        f64 = _dfsph_fields(g, o, sens)
        _log(f"1M: against the FP64 run of the same step: { {k: f'{v:.1e}' for k, v in f64.items()} }")
        assert f64["v"] < 1e-3 and f64["x"] < 1e-5, f64
    else:
        _log(f"1M: the FP64 run took {int(so['iterations'])}/{int(so['iterations_v'])} iterations")
    o.close()
    g.close()


def test_pcisph_spheres_full_size(ctx):
    t0 = time.time()
    spec = configs.pcisph_spheres()
    g = sphg.Scene(ctx, spec.params, spec.fluid, spec.boundary)
    g.step(20)
    o = po.Oracle(spec.params, spec.fluid, spec.boundary)
    _log(f"spheres: {spec.n} fluid, {spec.nb} boundary, 20 device steps ({time.time() - t0:.1f} s)")
    rho0 = spec.params.density0
    matched = None
    for attempt in range(3):
        _inject(g, o)
        g.step(1)
        o.step(1)
        gi, oi = g.scalars().iterations, int(o.scalars()["iterations"])
        assert abs(gi - oi) <= 1, (gi, oi)
        # the non-pressure forces precede the loop: compared on every step
        fa = rel_norm(g.get(sphg.FIELD_A), o.get(po.OF_ACC_STEP))
        fn = rel_norm(g.get(sphg.FIELD_NORMAL), o.get(po.OF_NORMAL))
        assert fa <= TOL and fn <= TOL, (fa, fn)
        if gi == oi:
            matched = {"a": fa, "normal": fn,
                       "drho": dev_err(g.get(sphg.FIELD_DRHO), o.get(po.OF_RHO), rho0, _floor(spec.params)),
                       "p": rel_norm(g.get(sphg.FIELD_P), o.get(po.OF_P)),
                       "pacc": rel_norm(g.get(sphg.FIELD_PACC), o.get(po.OF_PACC)),
                       "v": rel_norm(g.get(sphg.FIELD_V), o.get(po.OF_V)),
                       "x": rel_norm(g.get(sphg.FIELD_X), o.get(po.OF_X))}
            _log(f"spheres: {gi} iterations, errors { {k: f'{v:.1e}' for k, v in matched.items()} } "
                 f"({time.time() - t0:.1f} s)")
            break
    assert matched is not None
    _check(matched)
    o.close()
    g.close()


def test_wcsph_dam_32k_full_size(ctx):
    spec = configs.wcsph_dam_32k()
    g = sphg.Scene(ctx, spec.params, spec.fluid, spec.boundary)
    o = po.Oracle(spec.params, spec.fluid, spec.boundary)
    rho0 = spec.params.density0
    worst = {}
    for step in range(8):
        _inject(g, o)
        g.step(1)
        o.step(1)
        errs = {"drho": dev_err(g.get(sphg.FIELD_DRHO), o.get(po.OF_RHO), rho0, _floor(spec.params)),
                "p": rel_norm(g.get(sphg.FIELD_P), o.get(po.OF_P)),
                "a": rel_norm(g.get(sphg.FIELD_A), o.get(po.OF_ACC_STEP)),
                "v": rel_norm(g.get(sphg.FIELD_V), o.get(po.OF_V)),
                "x": rel_norm(g.get(sphg.FIELD_X), o.get(po.OF_X))}
        _check(errs)
        worst = {k: max(v, worst.get(k, 0.0)) for k, v in errs.items()}
    _log(f"wcsph 32k: 8 steps, worst errors { {k: f'{v:.1e}' for k, v in worst.items()} }")
    o.close()
    g.close()


def test_dense_lattice_build_reads_past_its_stage(ctx):
    """Half the rest spacing: ~260 neighbors per particle and 64-particle
    cells, so a wavefront's 9 candidate rows hold ~1,700 positions against a
    stage of 816; the rows past the stage are read from global memory."""
    from dataclasses import replace
    spec = configs.small_scene(sphg.DFSPH, n_side=16, spacing=0.05, jitter=0.01)
    r = spec.params.particle_radius
    fluid = (spec.fluid - spec.fluid.min(0)) * 0.5 + spec.fluid.min(0)
    params = replace(spec.params, max_neighbors=400)
    g = sphg.Scene(ctx, params, fluid, spec.boundary)
    o = po.Oracle(params, fluid, spec.boundary)
    dspec = configs.SceneSpec("dense", params, np.ascontiguousarray(fluid, dtype=np.float32), spec.boundary)
    npick, cmax = _grid_parity(g, o, dspec, sample=4096, dense=256)
    assert cmax > 200, cmax
    rho0 = params.density0
    assert dev_err(g.get(sphg.FIELD_DRHO), o.get(po.OF_RHO), rho0, _floor(params)) <= TOL
    assert rel_norm(g.get(sphg.FIELD_ALPHA), o.get(po.OF_ALPHA)) <= TOL
    _log(f"dense lattice: {spec.n} particles, longest list {cmax}, r {r}")
    o.close()
    g.close()


def test_batched_256_scenes(ctx):
    t0 = time.time()
    specs = configs.batched_dfsph()
    ntot = sum(s.n for s in specs)
    scenes = [sphg.Scene(ctx, s.params, s.fluid, s.boundary) for s in specs]
    ctx.synchronize()
    _log(f"batch: {len(specs)} scenes, {ntot} fluid particles resident ({time.time() - t0:.1f} s)")
    order = np.argsort([s.n for s in specs])
    small, large = int(order[0]), int(order[len(order) // 2])
    watch = {small: po.Oracle(specs[small].params, specs[small].fluid, specs[small].boundary),
             large: po.Oracle(specs[large].params, specs[large].fluid, specs[large].boundary)}
    b = sphg.Batch(ctx, scenes)
    compared = 0
    for step in range(5):
        starts = {k: _state(scenes[k]) for k in watch}
        for k, o in watch.items():
            _inject(scenes[k], o, starts[k])
        b.run(1)
        for k, o in watch.items():
            o.step(1)
            sg, so = scenes[k].scalars(), o.scalars()
            its = (sg.iterations, int(so["iterations"]), sg.iterations_v, int(so["iterations_v"]))
            assert abs(its[0] - its[1]) <= 1 and abs(its[2] - its[3]) <= 1, (k, its)
            if its[0] == its[1] and its[2] == its[3]:
                rho0 = specs[k].params.density0
                ref = _oracle_fields(o, rho0)
                sens = _eps_sensitive(o, starts[k], ref, rho0)
                errs = _dfsph_fields(scenes[k], o, sens, ref)
                errs["diverr_at_state"], _ = _diverr_at_state(scenes[k], o)
                _check(errs, {"diverr": TOL_DIVERR})
                compared += 1
    assert compared >= 8, compared
    b.close()
    _log(f"batch: 5 batched steps, {compared} scene-steps against the oracle ({time.time() - t0:.1f} s)")
    # the same scenes alone: bit-identical
    check = sorted({int(order[0]), int(order[-1]), int(order[len(order) // 2]), 17, 101, 200})
    xb = {k: (scenes[k].get(sphg.FIELD_X), scenes[k].get(sphg.FIELD_V)) for k in check}
    for s in scenes:
        s.close()
    for k in check:
        s = sphg.Scene(ctx, specs[k].params, specs[k].fluid, specs[k].boundary)
        s.step(5)
        assert np.array_equal(s.get(sphg.FIELD_X), xb[k][0]) and np.array_equal(s.get(sphg.FIELD_V), xb[k][1]), k
        s.close()
    for o in watch.values():
        o.close()
    _log(f"batch: {len(check)} scenes bit-identical alone ({time.time() - t0:.1f} s)")
