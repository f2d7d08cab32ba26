"""Known-answer solver cases of tests/golden/solver_closed_forms.json.

Each case builds its scene, runs it through a `runner` -- the FP64 oracle or
the HIP solver through the C ABI -- and compares the fields with the
decimal closed forms written by tests/golden/make_solver_closed_forms.py.
A failed comparison raises AssertionError naming the case and field.

A runner is called as runner(params, fluid, velocities, boundary, steps)
and returns (get, dts): get(name) gives a field after the last step
("x", "v", "a", "rho", "rhostar", "alpha", "diverr", "k", "kv", "p",
"pacc", "normal", "bvol", "iterations_v", "pcisph_delta"), dts the time step
of every step.
"""
import json
import os
from fractions import Fraction

import numpy as np

from sph_amd import sphg
from sph_amd.sphg import SceneParams

GOLD = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                   "solver_closed_forms.json")))


def _a(v):
    return np.array(v, dtype=np.float64) if not isinstance(v, str) else float(v)


def _f32_exact(a):
    a = np.asarray(a, dtype=np.float64)
    assert np.array_equal(a.astype(np.float32).astype(np.float64), a), "case inputs must be exact in FP32"
    return a


def _close(case, name, got, exp, tol):
    got = np.asarray(got, dtype=np.float64).reshape(np.shape(exp))
    exp = np.asarray(exp, dtype=np.float64)
    scale = np.abs(exp).max()
    err = np.abs(got - exp).max() / scale if scale > 0 else np.abs(got).max()
    assert err <= tol, f"{case}: {name} off by {err:.3e} (normwise relative), tolerance {tol:.0e}"
    return err


def _params(method, c, **kw):
    p = SceneParams(method=method, particle_radius=float(Fraction(c["r"])), gravity=(0.0, 0.0, 0.0))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _inputs(c):
    x = _f32_exact(_a(c["x"]))
    v = _f32_exact(_a(c["v"])) if "v" in c else np.zeros_like(x)
    b = _f32_exact(_a(c["boundary"]))
    return x, v, b


CASES = ["dfsph_pair_q0.1875", "dfsph_pair_q0.6875", "dfsph_cluster", "dfsph_dt_cfl", "dfsph_dt_gravity",
         "dfsph_dt_sqrt", "dfsph_dt_099", "dfsph_dt_1001_clamp", "wcsph_cluster", "pcisph_cluster",
         "surface_tension", "adhesion"]


def check(name, runner, tol, tol_dt=None):
    """Run case `name` and compare; returns {field: error}."""
    c = GOLD[name]
    e = c["expect"]
    x, v, b = _inputs(c) if "x" in c else (None, None, None)
    errs = {}
    if name.startswith("dfsph_pair"):
        dt = float(c["dt"])
        p = _params(sphg.DFSPH, c, time_step=dt, min_time_step=dt, max_time_step=dt, min_iterations=1,
                    max_iterations=1, min_iterations_v=0, eta_v=1e9)
        get, _ = runner(p, x, v, b, 1)
        for f in ("v", "k", "rhostar", "x"):
            errs[f] = _close(name, f, get(f), _a(e[f]), tol)
    elif name == "dfsph_cluster":
        dt = float(c["dt"])
        p = _params(sphg.DFSPH, c, time_step=dt, min_time_step=dt, max_time_step=dt, min_iterations=0,
                    max_iterations=1, eta=1e9, min_iterations_v=2, max_iterations_v=5, eta_v=1e9)
        get, _ = runner(p, x, v, b, 1)
        assert get("iterations_v") == 1, f"{name}: {get('iterations_v')} divergence iterations, expected 1"
        for f in ("x", "rho", "alpha", "kv", "v", "diverr"):
            errs[f] = _close(name, f, get(f), _a(e[f]), tol)
    elif name.startswith("dfsph_dt"):
        g = tuple(float(Fraction(t)) for t in c["g"])
        p = _params(sphg.DFSPH, c, gravity=g, time_step=float(c["dt0"]), min_time_step=float(c["dt_min"]),
                    max_time_step=float(c["dt_max"]), cfl_factor=float(c["cfl"]),
                    min_iterations=int(c["min_iterations"]), max_iterations=100)
        vv = _f32_exact(np.array([[float(Fraction(t)) for t in c["v"]]]))
        xx = _f32_exact(np.array([[0.5, 0.5, 0.5]]))
        far = np.array([[1.0, 1.0, 1.0]])
        _, dts = runner(p, xx, vv, far, int(c["steps"]))
        errs["dt"] = _close(name, "dt", dts, _a(e["dt"]), tol_dt or tol)
    elif name == "wcsph_cluster":
        dt = float(c["dt"])
        p = _params(sphg.WCSPH, c, time_step=dt, min_time_step=dt, max_time_step=dt, stiffness=1000.0, gamma=7.0,
                    viscosity_method=sphg.VISCOSITY_XSPH, viscosity=float(c["nu"]),
                    boundary_viscosity=float(c["nub"]))
        get, _ = runner(p, x, v, b, 1)
        for f in ("bvol", "rho", "p", "a"):
            errs[f] = _close(name, f, get(f), _a(e[f]), tol)
    elif name == "pcisph_cluster":
        dt = float(c["dt"])
        p = _params(sphg.PCISPH, c, time_step=dt, min_time_step=dt, max_time_step=dt, eta=0.01, min_iterations=1,
                    max_iterations=1)
        get, _ = runner(p, x, v, b, 1)
        if get("pcisph_delta") is not None:
            errs["delta"] = _close(name, "delta", get("pcisph_delta"), float(e["delta"]), tol)
        for f in ("rho", "p", "pacc", "v", "x"):
            errs[f] = _close(name, f, get(f), _a(e[f]), tol)
    elif name == "surface_tension":
        dt = float(c["dt"])
        p = _params(sphg.WCSPH, c, time_step=dt, min_time_step=dt, max_time_step=dt,
                    surface_tension_method=sphg.SURFACE_TENSION_AKINCI, surface_tension=float(c["gamma"]))
        get, _ = runner(p, x, v, b, 1)
        for f in ("rho", "normal", "a"):
            errs[f] = _close(name, f, get(f), _a(e[f]), tol)
    elif name == "adhesion":
        dt = float(c["dt"])
        p = _params(sphg.WCSPH, c, time_step=dt, min_time_step=dt, max_time_step=dt,
                    adhesion_method=sphg.ADHESION_AKINCI, adhesion=float(c["beta"]))
        get, _ = runner(p, x, v, b, 1)
        for f in ("bvol", "rho", "a"):
            errs[f] = _close(name, f, get(f), _a(e[f]), tol)
    else:
        raise KeyError(name)
    return errs


# Which cases each planted oracle mistake must fail (oracle/sph_oracle.cpp, enum Plant).
PLANT_CASES = {
    "KAPPA_X2": ["dfsph_pair_q0.1875", "dfsph_pair_q0.6875", "dfsph_cluster"],
    "KAPPA_DT": ["dfsph_pair_q0.1875", "dfsph_pair_q0.6875"],
    "KAPPA_RHO0": ["dfsph_pair_q0.1875", "dfsph_pair_q0.6875", "dfsph_cluster"],
    "K_NO_ALPHA": ["dfsph_pair_q0.1875", "dfsph_pair_q0.6875", "dfsph_cluster"],
    "SWEEP_SIGN": ["dfsph_pair_q0.1875", "dfsph_pair_q0.6875", "dfsph_cluster"],
    "KV_DT2": ["dfsph_cluster"],
    "DIV_LOOP_MAXV": ["dfsph_cluster"],
    "ALPHA_NO_SUMSQ": ["dfsph_pair_q0.1875", "dfsph_pair_q0.6875", "dfsph_cluster"],
    "DT_ABS_V": ["dfsph_dt_cfl", "dfsph_dt_gravity", "dfsph_dt_sqrt"],
    "DT_NO_G": ["dfsph_dt_gravity"],
    "DT_099_GE3": ["dfsph_dt_1001_clamp"],
    "WC_P_OVER_RHO": ["wcsph_cluster"],
    "WC_PJ_AS_PI": ["wcsph_cluster"],
    "WC_BOUNDARY_SYM": ["wcsph_cluster"],
    "XSPH_NO_DT": ["wcsph_cluster"],
    "XSPH_RHO_I": ["wcsph_cluster"],
    "PCI_NO_CLAMP": ["pcisph_cluster"],
    "PCI_CUBIC": ["pcisph_cluster"],
    "PCI_P_OVER_RHO": ["pcisph_cluster"],
    "ST_NPLUS": ["surface_tension"],
    "ST_K_HALF": ["surface_tension"],
    "ADH_COEF": ["adhesion"],
    "ADH_NO_RHO0": ["adhesion"],
}


def oracle_runner(params, x, v, b, steps):
    import oracle.pyoracle as po
    o = po.Oracle(params, x, b, velocities=v)
    dts = []
    for _ in range(steps):
        o.step(1)
        dts.append(o.scalars()["time_step"])
    fields = {"x": po.OF_X, "v": po.OF_V, "a": po.OF_ACC_STEP, "rho": po.OF_RHO, "rhostar": po.OF_RHOSTAR,
              "alpha": po.OF_ALPHA, "diverr": po.OF_DIVERR, "k": po.OF_K, "kv": po.OF_KV, "p": po.OF_P,
              "pacc": po.OF_PACC, "normal": po.OF_NORMAL, "bvol": po.OF_BVOL}
    sc = o.scalars()
    vals = {k: o.get(f) for k, f in fields.items()}
    vals["iterations_v"] = int(sc["iterations_v"])
    vals["pcisph_delta"] = sc["pcisph_delta"]
    o.close()
    return vals.__getitem__, dts


def gpu_runner(ctx):
    """The HIP solver through the C ABI (sph_amd.sphg) on context `ctx`."""
    def run(params, x, v, b, steps):
        s = sphg.Scene(ctx, params, x, b, velocities=v)
        dts = []
        for _ in range(steps):
            s.step(1)
            dts.append(s.scalars().time_step)
        fields = {"x": sphg.FIELD_X, "v": sphg.FIELD_V, "a": sphg.FIELD_A, "rho": sphg.FIELD_RHO,
                  "rhostar": sphg.FIELD_RHOSTAR, "alpha": sphg.FIELD_ALPHA, "diverr": sphg.FIELD_DIVERR,
                  "k": sphg.FIELD_K, "kv": sphg.FIELD_KV, "p": sphg.FIELD_P, "pacc": sphg.FIELD_PACC,
                  "normal": sphg.FIELD_NORMAL, "bvol": sphg.FIELD_BVOL}
        vals = {}
        for k, f in fields.items():
            try:
                vals[k] = s.get(f)
            except sphg.SphgError as e:     # a field this method or scene does not have
                assert e.code == sphg.EINVAL, e
        vals["iterations_v"] = int(s.scalars().iterations_v)
        vals["pcisph_delta"] = None          # a device scalar the ABI does not read back
        s.close()
        return vals.__getitem__, dts
    return run
