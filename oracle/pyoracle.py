"""ctypes wrapper of the FP64 CPU oracle (oracle/sph_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, never by the product path.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

OF_X, OF_V, OF_A, OF_RHO, OF_P, OF_ALPHA, OF_RHOSTAR = 0, 1, 2, 3, 4, 5, 6
OF_DIVERR, OF_K, OF_KV, OF_BVOL, OF_NORMAL, OF_PACC, OF_ACC_STEP, OF_BX = 7, 8, 9, 10, 11, 12, 13, 14
OF_PBN, OF_BDENS, OF_BPRESS, OF_PREDR = 15, 16, 17, 18
_VEC3 = {OF_X, OF_V, OF_A, OF_NORMAL, OF_PACC, OF_ACC_STEP, OF_PBN, OF_PREDR}

# Planted mistakes of the parity oracle (enum Plant in sph_oracle.cpp), for
# tests that show a check catches the mistake it is meant to catch.
PLANTS = ["NONE", "KAPPA_X2", "KAPPA_DT", "KAPPA_RHO0", "K_NO_ALPHA", "SWEEP_SIGN", "KV_DT2", "DIV_LOOP_MAXV",
          "ALPHA_NO_SUMSQ", "DT_ABS_V", "DT_NO_G", "DT_099_GE3", "WC_P_OVER_RHO", "WC_PJ_AS_PI", "WC_BOUNDARY_SYM",
          "XSPH_NO_DT", "XSPH_RHO_I", "PCI_NO_CLAMP", "PCI_CUBIC", "PCI_P_OVER_RHO", "ST_NPLUS", "ST_K_HALF",
          "ADH_COEF", "ADH_NO_RHO0"]


class OracleDesc(C.Structure):
    _fields_ = [
        ("method", C.c_int),
        ("particle_radius", C.c_double),
        ("density0", C.c_double),
        ("gravity", C.c_double * 3),
        ("time_step", C.c_double),
        ("min_time_step", C.c_double),
        ("max_time_step", C.c_double),
        ("fps", C.c_double),
        ("cfl_factor", C.c_double),
        ("stiffness", C.c_double),
        ("gamma", C.c_double),
        ("eta", C.c_double),
        ("eta_v", C.c_double),
        ("min_iterations", C.c_uint),
        ("max_iterations", C.c_uint),
        ("min_iterations_v", C.c_uint),
        ("max_iterations_v", C.c_uint),
        ("viscosity_method", C.c_int),
        ("viscosity", C.c_double),
        ("boundary_viscosity", C.c_double),
        ("surface_tension_method", C.c_int),
        ("surface_tension", C.c_double),
        ("adhesion_method", C.c_int),
        ("adhesion", C.c_double),
        ("cube_enabled", C.c_int),
        ("cube_min", C.c_double * 3),
        ("cube_max", C.c_double * 3),
        ("cube_normal_fct", C.c_double),
        ("cube_tangential_fct", C.c_double),
        ("boundary_method", C.c_int),
        ("pb_normal_fct", C.c_double),
        ("pb_tangential_fct", C.c_double),
    ]


class OracleEmitter(C.Structure):
    _fields_ = [
        ("type", C.c_uint),
        ("num_particles", C.c_uint),
        ("position", C.c_double * 3),
        ("velocity", C.c_double),
        ("rotation", C.c_double * 4),
        ("start_time", C.c_double),
        ("width", C.c_double),
        ("height", C.c_double),
        ("spacing", C.c_double),
    ]


_libs = {}


def load(fast: bool = False) -> C.CDLL:
    name = "liboracle_fast.so" if fast else "liboracle.so"
    if name in _libs:
        return _libs[name]
    path = os.path.join(_HERE, name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} missing: run `make -C oracle`")
    lib = C.CDLL(path)
    P, U, D = C.c_void_p, C.c_uint, C.c_double
    fp = C.POINTER(C.c_float)
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    up = C.POINTER(C.c_uint)
    sig = {
        "oracle_create": (P, [C.POINTER(OracleDesc)]),
        "oracle_destroy": (None, [P]),
        "oracle_set_threads": (C.c_int, [C.c_int]),
        "oracle_set_fluid": (C.c_int, [P, fp, fp, U]),
        "oracle_set_boundary": (C.c_int, [P, fp, U]),
        "oracle_init": (C.c_int, [P]),
        "oracle_plant": (C.c_int, [C.c_int]),
        "oracle_set_state": (C.c_int, [P, fp, fp, fp, fp, D, U]),
        "oracle_step": (C.c_int, [P, U]),
        "oracle_add_emitter": (C.c_int, [P, C.POINTER(OracleEmitter)]),
        "oracle_set_boundary_normals": (C.c_int, [P, fp, U]),
        "oracle_set_pb_normals": (C.c_int, [P, fp]),
        "oracle_sample_box_normals": (C.c_long, [dp, dp, D, C.c_int, dp, dp, C.c_long]),
        "oracle_round_state_f32": (C.c_int, [P]),
        "oracle_set_f32_positions": (C.c_int, [P, C.c_int]),
        "oracle_set_eps": (C.c_int, [P, D]),
        "oracle_divergence_error": (C.c_int, [P, fp]),
        "oracle_num_fluid": (U, [P]),
        "oracle_num_boundary": (U, [P]),
        "oracle_get": (C.c_int, [P, C.c_int, dp]),
        "oracle_scalars": (C.c_int, [P, dp]),
        "oracle_fluid_keys": (C.c_int, [P, ip]),
        "oracle_cell_points": (C.c_int, [P, U, up, U]),
        "oracle_candidates": (C.c_int, [P, U, up, U, C.c_int]),
        "oracle_sample_cube": (C.c_long, [dp, dp, D, dp, C.c_long]),
        "oracle_sample_sphere": (C.c_long, [dp, D, D, dp, C.c_long]),
        "oracle_fluid_block": (C.c_long, [dp, dp, D, dp, C.c_long]),
        "oracle_kernel_W": (D, [D, D]),
        "oracle_kernel_gradW": (None, [D, dp, dp]),
        "oracle_kernel_cohesion": (D, [D, D]),
        "oracle_kernel_adhesion": (D, [D, D]),
        "oracle_kernel_spiky_grad": (None, [D, dp, dp]),
        "oracle_kernel_poly6_grad": (None, [D, dp, dp]),
    }
    for n, (res, args) in sig.items():
        f = getattr(lib, n)
        f.restype = res
        f.argtypes = args
    _libs[name] = lib
    return lib


def plant(name: str) -> None:
    """Switch on one planted mistake ("NONE" restores the faithful oracle)
    in the parity build, for every oracle of the process."""
    n = load().oracle_plant(PLANTS.index(name))
    if n != len(PLANTS):
        raise RuntimeError(f"oracle plant table mismatch: library has {n} entries, wrapper {len(PLANTS)}")


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _fp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))


def make_desc(params) -> OracleDesc:
    """From sph_amd.sphg.SceneParams (same field names)."""
    d = OracleDesc()
    for name, _ in OracleDesc._fields_:
        v = getattr(params, name)
        if name in ("gravity", "cube_min", "cube_max"):
            getattr(d, name)[:] = [float(x) for x in v]
        else:
            setattr(d, name, v)
    return d


class Oracle:
    """FP64 restatement of the reference solver for one scene."""

    def __init__(self, params, fluid, boundary, velocities=None, init: bool = True, fast: bool = False,
                 threads: Optional[int] = None, boundary_normals=None):
        self.lib = load(fast)
        if threads:
            self.lib.oracle_set_threads(int(threads))
        self.params = params
        d = make_desc(params)
        self.h = self.lib.oracle_create(C.byref(d))
        x = np.ascontiguousarray(np.asarray(fluid, dtype=np.float32).reshape(-1, 3))
        v = None if velocities is None else np.ascontiguousarray(np.asarray(velocities, dtype=np.float32).reshape(-1, 3))
        b = np.ascontiguousarray(np.asarray(boundary, dtype=np.float32).reshape(-1, 3))
        self.n, self.nb = x.shape[0], b.shape[0]
        self.lib.oracle_set_fluid(self.h, _fp(x), _fp(v), self.n)
        self.lib.oracle_set_boundary(self.h, _fp(b), self.nb)
        if boundary_normals is not None:
            bn = np.ascontiguousarray(np.asarray(boundary_normals, dtype=np.float32).reshape(-1, 3))
            if self.lib.oracle_set_boundary_normals(self.h, _fp(bn), bn.shape[0]):
                raise ValueError("one boundary normal per boundary particle")
        if init:
            self.lib.oracle_init(self.h)

    def close(self):
        if self.h:
            self.lib.oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n: int = 1):
        self.lib.oracle_step(self.h, int(n))
        self.n = int(self.lib.oracle_num_fluid(self.h))

    def set_pb_normals(self, n):
        """The fluid's normal field of an injected state (PCISPH boundary method)."""
        a = np.ascontiguousarray(np.asarray(n, dtype=np.float32).reshape(-1, 3))
        self.lib.oracle_set_pb_normals(self.h, _fp(a))

    def add_emitter(self, e) -> int:
        """FluidModel::addEmitter: `e` has the fields of sph_amd.sphg.Emitter.
        Returns the particles per emitted group."""
        d = OracleEmitter()
        d.type, d.num_particles = int(e.type), int(e.num_particles)
        d.position[:] = [float(x) for x in e.position]
        d.velocity = float(e.velocity)
        d.rotation[:] = [float(x) for x in e.rotation]
        d.start_time, d.width, d.height, d.spacing = float(e.start_time), float(e.width), float(e.height), float(e.spacing)
        return int(self.lib.oracle_add_emitter(self.h, C.byref(d)))

    def set_state(self, x, v, k=None, kv=None, time_step=None, last_iterations=0):
        x = np.ascontiguousarray(np.asarray(x, dtype=np.float32).reshape(-1, 3))
        v = np.ascontiguousarray(np.asarray(v, dtype=np.float32).reshape(-1, 3))
        k = None if k is None else np.ascontiguousarray(np.asarray(k, dtype=np.float32))
        kv = None if kv is None else np.ascontiguousarray(np.asarray(kv, dtype=np.float32))
        dt = self.params.time_step if time_step is None else float(time_step)
        self.lib.oracle_set_state(self.h, _fp(x), _fp(v), _fp(k), _fp(kv), dt, int(last_iterations))

    def round_state_f32(self):
        """Round x, v, kappa and kappa_v to FP32, keeping the neighbor lists,
        densities and alpha factors (the state a solver that stores particle
        fields in FP32 carries into its next step)."""
        self.lib.oracle_round_state_f32(self.h)

    def set_f32_positions(self, on=True):
        """Round DFSPH's updated positions to FP32 before the neighbor
        search, densities and alpha factors of the step."""
        self.lib.oracle_set_f32_positions(self.h, 1 if on else 0)

    def set_eps(self, eps: float = 1e-5):
        """The threshold of the DFSPH eps branches (1e-5 in the reference)."""
        self.lib.oracle_set_eps(self.h, float(eps))

    def divergence_error(self, velocities) -> np.ndarray:
        """The divergence error (src/DFSPHSolver.cpp:408-492) of the current
        positions and lists with `velocities` (n x 3) in place of the
        oracle's own; also left in OF_DIVERR."""
        v = np.ascontiguousarray(velocities, dtype=np.float32).reshape(-1, 3)
        assert v.shape[0] == self.n
        self.lib.oracle_divergence_error(self.h, _fp(v))
        return self.get(OF_DIVERR)

    def get(self, field: int) -> np.ndarray:
        if field in (OF_BVOL, OF_BDENS, OF_BPRESS):
            out = np.zeros(max(self.nb, 1))
            self.lib.oracle_get(self.h, field, _dp(out))
            return out[: self.nb]
        if field == OF_BX:
            out = np.zeros((max(self.nb, 1), 3))
            self.lib.oracle_get(self.h, field, _dp(out))
            return out[: self.nb]
        out = np.zeros((self.n, 3)) if field in _VEC3 else np.zeros(self.n)
        self.lib.oracle_get(self.h, field, _dp(out))
        return out

    def scalars(self) -> dict:
        out = np.zeros(16)
        self.lib.oracle_scalars(self.h, _dp(out))
        keys = ["time", "time_step", "iterations", "iterations_v", "avg_density_error", "avg_divergence_error",
                "max_density_error", "steps", "sum_iterations", "sum_iterations_v", "max_vel", "pcisph_delta",
                "cfl_time_step", "avg_density_error0", "avg_divergence_error0"]
        return {k: float(out[i]) for i, k in enumerate(keys)}

    def keys(self) -> np.ndarray:
        out = np.zeros((self.n, 3), dtype=np.int32)
        self.lib.oracle_fluid_keys(self.h, out.ctypes.data_as(C.POINTER(C.c_int)))
        return out

    def cell_points(self, i: int) -> np.ndarray:
        cap = 4096
        out = np.zeros(cap, dtype=np.uint32)
        n = self.lib.oracle_cell_points(self.h, int(i), out.ctypes.data_as(C.POINTER(C.c_uint)), cap)
        return out[:n].copy()

    def candidates(self, i: int, boundary: bool = False) -> np.ndarray:
        cap = 1 << 16
        out = np.zeros(cap, dtype=np.uint32)
        n = self.lib.oracle_candidates(self.h, int(i), out.ctypes.data_as(C.POINTER(C.c_uint)), cap,
                                       1 if boundary else 0)
        return out[:n].copy()


def _sampler(fn, *args):
    n = fn(*args, None, 0)
    out = np.zeros((max(n, 1), 3))
    fn(*args, _dp(out), n)
    return out[:n]


def sample_cube(mn, mx, support_radius):
    lib = load()
    return _sampler(lib.oracle_sample_cube, _dp(np.asarray(mn, float)), _dp(np.asarray(mx, float)),
                    float(support_radius))


def sample_sphere(center, radius, particle_radius):
    lib = load()
    return _sampler(lib.oracle_sample_sphere, _dp(np.asarray(center, float)), float(radius), float(particle_radius))


def sample_box_normals(mn, mx, support_radius, inverted):
    """PCISPHBoundaryModel::sampleCube + addCube: points and their normals."""
    lib = load()
    a, b = _dp(np.asarray(mn, float)), _dp(np.asarray(mx, float))
    n = lib.oracle_sample_box_normals(a, b, float(support_radius), int(bool(inverted)), None, None, 0)
    pts, nrm = np.zeros((max(n, 1), 3)), np.zeros((max(n, 1), 3))
    lib.oracle_sample_box_normals(a, b, float(support_radius), int(bool(inverted)), _dp(pts), _dp(nrm), n)
    return pts[:n], nrm[:n]


def sphere_normals(points, center, inverted):
    """PCISPHBoundaryModel::addSphere (src/PCISPHBoundaryModel.cpp:30-52):
    n = inv (center - point), inv = 1 if inverted else -1, unnormalised."""
    inv = 1.0 if inverted else -1.0
    return inv * (np.asarray(center, float)[None, :] - np.asarray(points, float))


def fluid_block(mn, mx, particle_radius):
    lib = load()
    return _sampler(lib.oracle_fluid_block, _dp(np.asarray(mn, float)), _dp(np.asarray(mx, float)),
                    float(particle_radius))


def kernel_W(h, r):
    return load().oracle_kernel_W(float(h), float(r))


def kernel_gradW(h, r):
    out = np.zeros(3)
    load().oracle_kernel_gradW(float(h), _dp(np.asarray(r, float)), _dp(out))
    return out


def kernel_spiky_grad(h, r):
    out = np.zeros(3)
    load().oracle_kernel_spiky_grad(float(h), _dp(np.asarray(r, float)), _dp(out))
    return out


def kernel_poly6_grad(h, r):
    out = np.zeros(3)
    load().oracle_kernel_poly6_grad(float(h), _dp(np.asarray(r, float)), _dp(out))
    return out


def kernel_cohesion(h, r):
    return load().oracle_kernel_cohesion(float(h), float(r))


def kernel_adhesion(h, r):
    return load().oracle_kernel_adhesion(float(h), float(r))
