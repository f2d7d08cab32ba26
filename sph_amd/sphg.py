"""ctypes binding of libsphg (include/sphg.h), the gfx950 SPH solver.

The library is loaded from this package directory; a missing or unloadable
library raises immediately -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field, fields
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsphg.so")

WCSPH, PCISPH, DFSPH = 0, 1, 2
# status codes (SPHG_OK, SPHG_E*)
OK, EINVAL, EHIP, ENOMEM, ESTATE, EOVERFLOW = 0, -1, -2, -3, -4, -5
BOUNDARY_CUBE, BOUNDARY_PCISPH, BOUNDARY_AKINCI = 0, 1, 2
VISCOSITY_STANDARD, VISCOSITY_ARTIFICIAL, VISCOSITY_XSPH = 0, 1, 2
SURFACE_TENSION_AKINCI, ADHESION_AKINCI = 3, 4

FIELD_X, FIELD_V, FIELD_A, FIELD_RHO, FIELD_P = 0, 1, 2, 3, 4
FIELD_ALPHA, FIELD_RHOSTAR, FIELD_DIVERR, FIELD_K, FIELD_KV = 5, 6, 7, 8, 9
FIELD_BVOL, FIELD_NORMAL, FIELD_PACC, FIELD_KEYS, FIELD_NNB, FIELD_NGRAD = 10, 11, 12, 13, 14, 15
FIELD_PBNORMAL, FIELD_BDENSITY, FIELD_BPRESSURE = 16, 17, 18
FIELD_DRHO, FIELD_DRHOSTAR, FIELD_NSH = 19, 20, 21

_VEC3 = {FIELD_X, FIELD_V, FIELD_A, FIELD_NORMAL, FIELD_PACC, FIELD_PBNORMAL}
_INT = {FIELD_KEYS, FIELD_NNB, FIELD_NGRAD, FIELD_NSH}


class SphgError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"sphg error {code}: {msg}")
        self.code = code


class SceneDesc(C.Structure):
    _fields_ = [
        ("method", C.c_int32),
        ("boundary_method", C.c_int32),
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
        ("min_iterations", C.c_uint32),
        ("max_iterations", C.c_uint32),
        ("min_iterations_v", C.c_uint32),
        ("max_iterations_v", C.c_uint32),
        ("viscosity_method", C.c_int32),
        ("viscosity", C.c_double),
        ("boundary_viscosity", C.c_double),
        ("surface_tension_method", C.c_int32),
        ("surface_tension", C.c_double),
        ("adhesion_method", C.c_int32),
        ("adhesion", C.c_double),
        ("max_neighbors", C.c_uint32),
        ("domain_min", C.c_double * 3),
        ("domain_max", C.c_double * 3),
        ("cube_enabled", C.c_int32),
        ("cube_min", C.c_double * 3),
        ("cube_max", C.c_double * 3),
        ("cube_normal_fct", C.c_double),
        ("cube_tangential_fct", C.c_double),
        ("pb_normal_fct", C.c_double),
        ("pb_tangential_fct", C.c_double),
    ]


class SceneScalars(C.Structure):
    _fields_ = [
        ("time", C.c_double),
        ("time_step", C.c_double),
        ("max_vel", C.c_double),
        ("cfl_time_step", C.c_double),
        ("avg_density_error", C.c_double),
        ("avg_density_error0", C.c_double),
        ("avg_divergence_error", C.c_double),
        ("avg_divergence_error0", C.c_double),
        ("max_density_error", C.c_double),
        ("iterations", C.c_int32),
        ("iterations_v", C.c_int32),
        ("steps", C.c_uint32),
        ("sum_iterations", C.c_uint32),
        ("sum_iterations_v", C.c_uint32),
        ("neighbor_overflow", C.c_uint32),
        ("pcisph_shell_steps", C.c_uint32),
        ("pcisph_walk_steps", C.c_uint32),
    ]


class StepStats(C.Structure):
    _fields_ = [
        ("gpu_ms", C.c_double),
        ("steps", C.c_uint32),
        ("iterations", C.c_uint32),
        ("iterations_v", C.c_uint32),
        ("host_syncs", C.c_uint32),
    ]


class TuneResult(C.Structure):
    """The launch schedule an autotune kept (sphg_tune_result)."""
    _fields_ = [
        ("density_slack", C.c_int32),
        ("divergence_slack", C.c_int32),
        ("replay", C.c_int32),
        ("candidates", C.c_uint32),
        ("steps", C.c_uint32),
        ("ms_per_step", C.c_double),
        ("default_ms_per_step", C.c_double),
    ]


class EmitterDesc(C.Structure):
    _fields_ = [
        ("type", C.c_uint32),
        ("num_particles", C.c_uint32),
        ("position", C.c_double * 3),
        ("velocity", C.c_double),
        ("rotation", C.c_double * 4),
        ("start_time", C.c_double),
        ("width", C.c_double),
        ("height", C.c_double),
        ("spacing", C.c_double),
    ]


EMITTER_SQUARE, EMITTER_CIRCLE = 0, 1


@dataclass
class Emitter:
    """A particle emitter with the reference scene file's emitter keys
    (src/SceneLoader.cpp:123-139): a square (width x height) or circle
    (radius width) group at spacing 2r, emitted along the local +z axis at
    `velocity`, rotated by the quaternion `rotation` = (w, x, y, z)."""

    type: int = EMITTER_SQUARE
    num_particles: int = 0
    position: Sequence[float] = (0.0, 0.0, 0.0)
    velocity: float = 1.0
    rotation: Sequence[float] = (1.0, 0.0, 0.0, 0.0)
    start_time: float = 0.0
    width: float = 0.1
    height: float = 0.1
    spacing: float = 1.0

    def to_desc(self) -> EmitterDesc:
        d = EmitterDesc()
        d.type, d.num_particles = int(self.type), int(self.num_particles)
        d.position[:] = [float(x) for x in self.position]
        d.velocity = float(self.velocity)
        d.rotation[:] = [float(x) for x in self.rotation]
        d.start_time, d.width, d.height, d.spacing = (float(self.start_time), float(self.width), float(self.height),
                                                      float(self.spacing))
        return d


@dataclass
class SceneParams:
    """Scene parameters with the reference scene-JSON meaning
    (src/SceneLoader.cpp:38-141 of the reference)."""

    method: int = DFSPH
    particle_radius: float = 0.025
    density0: float = 1000.0
    gravity: Sequence[float] = (0.0, 0.0, -9.81)
    time_step: float = 1e-3
    min_time_step: float = 1e-6
    max_time_step: float = 5e-3
    fps: float = 25.0
    cfl_factor: float = 0.4
    stiffness: float = 1000.0
    gamma: float = 7.0
    eta: float = 1e-4
    eta_v: float = 1e-3
    min_iterations: int = 2
    max_iterations: int = 100
    min_iterations_v: int = 1
    max_iterations_v: int = 100
    viscosity_method: int = -1
    viscosity: float = 0.0
    boundary_viscosity: float = 0.0
    surface_tension_method: int = -1
    surface_tension: float = 0.0
    adhesion_method: int = -1
    adhesion: float = 0.0
    max_neighbors: int = 0
    domain_min: Sequence[float] = (0.0, 0.0, 0.0)
    domain_max: Sequence[float] = (0.0, 0.0, 0.0)
    boundary_method: int = 2          # BOUNDARY_AKINCI; BOUNDARY_CUBE = no boundary particles
    cube_enabled: int = 0             # CubeBoundaryModel clamp (boundary_method 0)
    cube_min: Sequence[float] = (0.0, 0.0, 0.0)
    cube_max: Sequence[float] = (0.0, 0.0, 0.0)
    cube_normal_fct: float = 0.0
    cube_tangential_fct: float = 1.0
    pb_normal_fct: float = 0.0        # PCISPHBoundaryModel (boundary_method 1)
    pb_tangential_fct: float = 1.0

    @property
    def support_radius(self) -> float:
        return 4.0 * self.particle_radius

    def to_desc(self) -> SceneDesc:
        d = SceneDesc()
        for f in fields(self):
            v = getattr(self, f.name)
            if f.name in ("gravity", "domain_min", "domain_max", "cube_min", "cube_max"):
                getattr(d, f.name)[:] = [float(x) for x in v]
            else:
                setattr(d, f.name, v)
        return d


_lib: Optional[C.CDLL] = None


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load libsphg (once).  Raises if it is missing: the solver has no CPU path."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `make -C sph_amd/csrc` or __graft_entry__.build()")
    lib = C.CDLL(path)
    P, I32, U32, U64, I64, D = C.c_void_p, C.c_int32, C.c_uint32, C.c_uint64, C.c_int64, C.c_double
    fp = C.POINTER(C.c_float)
    dp = C.POINTER(C.c_double)
    sig = {
        "sphg_version": (C.c_char_p, []),
        "sphg_last_error": (C.c_char_p, []),
        "sphg_ctx_create": (C.c_int, [I32, C.POINTER(P)]),
        "sphg_ctx_destroy": (C.c_int, [P]),
        "sphg_ctx_synchronize": (C.c_int, [P]),
        "sphg_ctx_graphs": (C.c_int, [P, C.c_int32]),
        "sphg_scene_create": (C.c_int, [P, C.POINTER(SceneDesc), C.POINTER(P)]),
        "sphg_scene_destroy": (C.c_int, [P]),
        "sphg_scene_set_fluid": (C.c_int, [P, fp, fp, U32]),
        "sphg_scene_set_boundary": (C.c_int, [P, fp, U32]),
        "sphg_scene_add_emitter": (C.c_int, [P, C.POINTER(EmitterDesc)]),
        "sphg_scene_set_boundary_normals": (C.c_int, [P, fp, U32]),
        "sphg_scene_set_pb_normals": (C.c_int, [P, fp]),
        "sphg_sample_box_normals": (I64, [dp, dp, D, I32, fp, fp, I64]),
        "sphg_sample_sphere_normals": (I64, [dp, D, D, I32, fp, fp, I64]),
        "sphg_scene_capacity": (C.c_int, [P, C.POINTER(U32)]),
        "sphg_snapshot_create": (C.c_int, [P, C.POINTER(P)]),
        "sphg_snapshot_begin": (C.c_int, [P]),
        "sphg_snapshot_wait": (C.c_int, [P, C.POINTER(fp), C.POINTER(fp), C.POINTER(U32), dp, dp]),
        "sphg_snapshot_destroy": (C.c_int, [P]),
        "sphg_scene_init": (C.c_int, [P]),
        "sphg_scene_set_state": (C.c_int, [P, fp, fp, fp, fp, D, U32]),
        "sphg_scene_step": (C.c_int, [P, U32, C.POINTER(StepStats)]),
        "sphg_scene_get": (C.c_int, [P, I32, P]),
        "sphg_scene_get_scalars": (C.c_int, [P, C.POINTER(SceneScalars)]),
        "sphg_scene_get_sorted": (C.c_int, [P, C.POINTER(C.c_uint32), C.POINTER(C.c_int32)]),
        "sphg_scene_get_neighbors": (C.c_int, [P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), U32]),
        "sphg_scene_info": (C.c_int, [P, C.POINTER(U32), C.POINTER(U32), C.POINTER(U32), C.POINTER(U64)]),
        "sphg_batch_step": (C.c_int, [P, C.POINTER(P), U32, U32, C.POINTER(StepStats)]),
        "sphg_batch_step_noalloc": (C.c_int, [P, C.POINTER(P), U32, U32, C.POINTER(StepStats)]),
        "sphg_batch_create": (C.c_int, [P, C.POINTER(P), U32, C.POINTER(P)]),
        "sphg_batch_run": (C.c_int, [P, U32, C.POINTER(StepStats)]),
        "sphg_batch_destroy": (C.c_int, [P]),
        "sphg_scene_autotune": (C.c_int, [P, U32, U32, C.POINTER(TuneResult)]),
        "sphg_batch_autotune": (C.c_int, [P, U32, U32, C.POINTER(TuneResult)]),
        "sphg_scene_plan_bytes": (C.c_int, [P, C.POINTER(U64)]),
        "sphg_ctx_kernel_timing": (C.c_int, [P, I32]),
        "sphg_ctx_kernel_time": (C.c_int, [P, C.c_char_p, C.POINTER(U64), dp]),
        "sphg_ctx_kernel_timing_only": (C.c_int, [P, C.POINTER(C.c_char_p), U32]),
        "sphg_ctx_kernel_timing_stride": (C.c_int, [P, U32]),
        "sphg_sample_box": (I64, [dp, dp, D, fp, I64]),
        "sphg_sample_sphere": (I64, [dp, D, D, fp, I64]),
        "sphg_fluid_block": (I64, [dp, dp, D, D, U64, fp, I64]),
        "sphg_fluid_spheres": (I64, [dp, dp, D, dp, dp, I32, fp, I64]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _check(rc: int) -> None:
    if rc != 0:
        raise SphgError(rc, load().sphg_last_error().decode())


def _fptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f32(a, shape_last=None) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    return a


# ----------------------------------------------------------------- geometry


def _sample(fn, *args) -> np.ndarray:
    n = fn(*args, None, 0)
    if n < 0:
        raise ValueError("invalid sampler arguments")
    out = np.zeros((max(n, 1), 3), dtype=np.float32)
    fn(*args, _fptr(out), n)
    return out[:n]


def sample_box(mn, mx, support_radius: float) -> np.ndarray:
    """AkinciBoundaryModel::sampleCube (src/AkinciBoundayModel.cpp:100-134)."""
    lib = load()
    a = np.asarray(mn, dtype=np.float64)
    b = np.asarray(mx, dtype=np.float64)
    return _sample(lib.sphg_sample_box, _dptr(a), _dptr(b), float(support_radius))


def sample_sphere(center, radius: float, particle_radius: float) -> np.ndarray:
    """AkinciBoundaryModel::sampleSphere (src/AkinciBoundayModel.cpp:136-164)."""
    lib = load()
    c = np.asarray(center, dtype=np.float64)
    return _sample(lib.sphg_sample_sphere, _dptr(c), float(radius), float(particle_radius))


def _sample_normals(fn, *args):
    n = fn(*args, None, None, 0)
    if n < 0:
        raise SphgError("invalid sampler arguments")
    pts = np.zeros((max(n, 1), 3), dtype=np.float32)
    nrm = np.zeros((max(n, 1), 3), dtype=np.float32)
    fn(*args, _fptr(pts), _fptr(nrm), n)
    return pts[:n], nrm[:n]


def sample_box_normals(mn, mx, support_radius: float, inverted: bool):
    """PCISPHBoundaryModel::addCube (src/PCISPHBoundaryModel.cpp:7-27,60-189):
    points and normals (inward when inverted)."""
    lib = load()
    a = np.asarray(mn, dtype=np.float64)
    b = np.asarray(mx, dtype=np.float64)
    return _sample_normals(lib.sphg_sample_box_normals, _dptr(a), _dptr(b), float(support_radius), int(bool(inverted)))


def sample_sphere_normals(center, radius: float, particle_radius: float, inverted: bool):
    """PCISPHBoundaryModel::addSphere (src/PCISPHBoundaryModel.cpp:29-52)."""
    lib = load()
    c = np.asarray(center, dtype=np.float64)
    return _sample_normals(lib.sphg_sample_sphere_normals, _dptr(c), float(radius), float(particle_radius),
                           int(bool(inverted)))


def fluid_block(mn, mx, particle_radius: float, jitter: float = 0.0, seed: int = 0) -> np.ndarray:
    """Simulation::buildFluidBlock (src/Simulation.cpp:512-539) + seeded jitter."""
    lib = load()
    a = np.asarray(mn, dtype=np.float64)
    b = np.asarray(mx, dtype=np.float64)
    return _sample(lib.sphg_fluid_block, _dptr(a), _dptr(b), float(particle_radius), float(jitter), int(seed))


def fluid_spheres(mn, mx, particle_radius: float, centers, radii) -> np.ndarray:
    lib = load()
    a = np.asarray(mn, dtype=np.float64)
    b = np.asarray(mx, dtype=np.float64)
    c = np.ascontiguousarray(np.asarray(centers, dtype=np.float64).reshape(-1, 3))
    r = np.ascontiguousarray(np.asarray(radii, dtype=np.float64).reshape(-1))
    return _sample(lib.sphg_fluid_spheres, _dptr(a), _dptr(b), float(particle_radius), _dptr(c), _dptr(r),
                   int(len(r)))


# ------------------------------------------------------------- contexts


class Context:
    """One device, one stream (sphg_ctx)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        self.device = int(device)
        h = C.c_void_p()
        _check(self.lib.sphg_ctx_create(int(device), C.byref(h)))
        self.h = h
        self._scenes = []

    def close(self) -> None:
        if self.h:
            for s in self._scenes:
                s.h = None
            self.lib.sphg_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self) -> None:
        _check(self.lib.sphg_ctx_synchronize(self.h))

    def set_graphs(self, enable: bool) -> None:
        """Replay captured step segments, or launch every kernel one by one
        (default); results are identical either way."""
        _check(self.lib.sphg_ctx_graphs(self.h, 1 if enable else 0))

    def scene(self, params: SceneParams, fluid, boundary, velocities=None) -> "Scene":
        return Scene(self, params, fluid, boundary, velocities)

    def batch_step(self, scenes: Sequence["Scene"], nsteps: int = 1) -> StepStats:
        arr = (C.c_void_p * len(scenes))(*[s.h for s in scenes])
        st = StepStats()
        _check(self.lib.sphg_batch_step(self.h, arr, len(scenes), int(nsteps), C.byref(st)))
        return st

    def batch_step_noalloc(self, scenes: Sequence["Scene"], nsteps: int = 1) -> StepStats:
        """batch_step with the batch's tables in the library's static device
        array: no device memory is allocated (sphg_batch_step_noalloc)."""
        arr = (C.c_void_p * len(scenes))(*[s.h for s in scenes])
        st = StepStats()
        _check(self.lib.sphg_batch_step_noalloc(self.h, arr, len(scenes), int(nsteps), C.byref(st)))
        return st

    def kernel_timing(self, enable: bool, only: Optional[Sequence[str]] = None) -> None:
        """Start (clearing the totals) or stop per-kernel event timing; with
        `only`, record just the kernels named."""
        if enable and only is not None:
            names = (C.c_char_p * max(len(only), 1))(*[n.encode() for n in only])
            _check(self.lib.sphg_ctx_kernel_timing_only(self.h, names, len(only)))
        else:
            _check(self.lib.sphg_ctx_kernel_timing(self.h, 1 if enable else 0))

    def kernel_timing_stride(self, stride: int) -> None:
        """While timing, record every `stride`-th launch of each kernel only
        (a sample; enabling timing resets it to 1)."""
        _check(self.lib.sphg_ctx_kernel_timing_stride(self.h, int(stride)))

    def kernel_time(self, name: str):
        n = C.c_uint64()
        ms = C.c_double()
        _check(self.lib.sphg_ctx_kernel_time(self.h, name.encode(), C.byref(n), C.byref(ms)))
        return int(n.value), float(ms.value)


class Batch:
    """Scenes of one method stepped together, kept between calls (sphg_batch).

    Close it before closing any of its scenes."""

    def __init__(self, ctx: "Context", scenes: Sequence["Scene"]):
        self.ctx = ctx
        self.lib = ctx.lib
        self.scenes = list(scenes)
        arr = (C.c_void_p * len(self.scenes))(*[s.h for s in self.scenes])
        h = C.c_void_p()
        _check(self.lib.sphg_batch_create(ctx.h, arr, len(self.scenes), C.byref(h)))
        self.h = h

    def run(self, nsteps: int = 1) -> StepStats:
        st = StepStats()
        _check(self.lib.sphg_batch_run(self.h, int(nsteps), C.byref(st)))
        return st

    def autotune(self, steps_per_trial: int = 5, rounds: int = 3) -> TuneResult:
        """Pick the fastest launch schedule for this batch; the trials are
        real steps (every scene advances by TuneResult.steps)."""
        r = TuneResult()
        _check(self.lib.sphg_batch_autotune(self.h, int(steps_per_trial), int(rounds), C.byref(r)))
        return r

    def close(self) -> None:
        if self.h:
            self.lib.sphg_batch_destroy(self.h)
            self.h = None


class Snapshot:
    """Asynchronous frame of a scene (sphg_snapshot): begin() enqueues the
    read-back of the active particles' positions and velocities (id order)
    and returns at once; wait() returns them with the time and time step of
    the moment begin() was called."""

    def __init__(self, scene: "Scene"):
        self.scene = scene
        self.lib = scene.lib
        h = C.c_void_p()
        _check(self.lib.sphg_snapshot_create(scene.h, C.byref(h)))
        self.h = h

    def begin(self) -> "Snapshot":
        _check(self.lib.sphg_snapshot_begin(self.h))
        return self

    def wait(self):
        x, v = C.POINTER(C.c_float)(), C.POINTER(C.c_float)()
        n = C.c_uint32()
        t, dt = C.c_double(), C.c_double()
        _check(self.lib.sphg_snapshot_wait(self.h, C.byref(x), C.byref(v), C.byref(n), C.byref(t), C.byref(dt)))
        k = int(n.value)
        xs = np.ctypeslib.as_array(x, shape=(k, 3)).copy() if k else np.zeros((0, 3), np.float32)
        vs = np.ctypeslib.as_array(v, shape=(k, 3)).copy() if k else np.zeros((0, 3), np.float32)
        return xs, vs, float(t.value), float(dt.value)

    def close(self) -> None:
        if self.h:
            self.lib.sphg_snapshot_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    """One fluid + Akinci boundary scene resident on the device (sphg_scene)."""

    def __init__(self, ctx: Context, params: SceneParams, fluid, boundary, velocities=None, init: bool = True,
                 emitters: Sequence[Emitter] = (), boundary_normals=None):
        self.ctx = ctx
        self.lib = ctx.lib
        self.params = params
        h = C.c_void_p()
        desc = params.to_desc()
        _check(self.lib.sphg_scene_create(ctx.h, C.byref(desc), C.byref(h)))
        self.h = h
        ctx._scenes.append(self)
        x = _f32(fluid).reshape(-1, 3)
        v = None if velocities is None else _f32(velocities).reshape(-1, 3)
        b = _f32(boundary).reshape(-1, 3)
        self.nb = b.shape[0]
        self.emitters = list(emitters)
        if x.shape[0]:
            _check(self.lib.sphg_scene_set_fluid(self.h, _fptr(x), _fptr(v), x.shape[0]))
        _check(self.lib.sphg_scene_set_boundary(self.h, _fptr(b), self.nb))
        if boundary_normals is not None:
            bn = _f32(boundary_normals).reshape(-1, 3)
            _check(self.lib.sphg_scene_set_boundary_normals(self.h, _fptr(bn), bn.shape[0]))
        for e in self.emitters:
            d = e.to_desc()
            _check(self.lib.sphg_scene_add_emitter(self.h, C.byref(d)))
        if init:
            _check(self.lib.sphg_scene_init(self.h))

    @property
    def n(self) -> int:
        """Active fluid particles (FluidModel::getNumActiveParticles)."""
        v = C.c_uint32()
        _check(self.lib.sphg_scene_info(self.h, C.byref(v), None, None, None))
        return int(v.value)

    @property
    def capacity(self) -> int:
        """Fluid particles plus every emitter's budget (FluidModel::getNumParticles)."""
        v = C.c_uint32()
        _check(self.lib.sphg_scene_capacity(self.h, C.byref(v)))
        return int(v.value)

    def init(self) -> None:
        """Allocate and initialise on the device (when created with init=False)."""
        _check(self.lib.sphg_scene_init(self.h))

    def plan_bytes(self) -> int:
        """Device bytes the scene holds once initialised (known before init)."""
        b = C.c_uint64()
        _check(self.lib.sphg_scene_plan_bytes(self.h, C.byref(b)))
        return int(b.value)

    def close(self) -> None:
        if self.h:
            self.lib.sphg_scene_destroy(self.h)
            self.h = None
            if self in self.ctx._scenes:
                self.ctx._scenes.remove(self)

    def step(self, nsteps: int = 1) -> StepStats:
        st = StepStats()
        _check(self.lib.sphg_scene_step(self.h, int(nsteps), C.byref(st)))
        return st

    def autotune(self, steps_per_trial: int = 5, rounds: int = 3) -> TuneResult:
        """Pick the fastest launch schedule for this scene's stepping; the
        trials are real steps (the scene advances by TuneResult.steps)."""
        r = TuneResult()
        _check(self.lib.sphg_scene_autotune(self.h, int(steps_per_trial), int(rounds), C.byref(r)))
        return r

    def set_state(self, x, v, k=None, kv=None, time_step: float = None, last_iterations: int = 0) -> None:
        x = _f32(x).reshape(-1, 3)
        v = _f32(v).reshape(-1, 3)
        k = None if k is None else _f32(k).reshape(-1)
        kv = None if kv is None else _f32(kv).reshape(-1)
        dt = self.params.time_step if time_step is None else float(time_step)
        _check(self.lib.sphg_scene_set_state(self.h, _fptr(x), _fptr(v), _fptr(k), _fptr(kv), dt,
                                             int(last_iterations)))

    def set_pb_normals(self, n) -> None:
        """The fluid normal field of the PCISPH boundary method (id order)."""
        a = _f32(n).reshape(-1, 3)
        _check(self.lib.sphg_scene_set_pb_normals(self.h, _fptr(a)))

    def get(self, fld: int) -> np.ndarray:
        if fld in (FIELD_BVOL, FIELD_BDENSITY, FIELD_BPRESSURE):
            out = np.zeros(max(self.nb, 1), dtype=np.float32)
            _check(self.lib.sphg_scene_get(self.h, fld, out.ctypes.data_as(C.c_void_p)))
            return out[: self.nb]
        if fld in _VEC3 or fld == FIELD_KEYS:
            dt = np.int32 if fld == FIELD_KEYS else np.float32
            out = np.zeros((self.n, 3), dtype=dt)
        else:
            out = np.zeros(self.n, dtype=np.int32 if fld in _INT else np.float32)
        _check(self.lib.sphg_scene_get(self.h, fld, out.ctypes.data_as(C.c_void_p)))
        return out

    def snapshot(self) -> Snapshot:
        """A reusable asynchronous frame object for this scene."""
        return Snapshot(self)

    def scalars(self) -> SceneScalars:
        s = SceneScalars()
        _check(self.lib.sphg_scene_get_scalars(self.h, C.byref(s)))
        return s

    def sorted_ids(self):
        ids = np.zeros(self.n, dtype=np.uint32)
        cells = np.zeros(self.n, dtype=np.int32)
        _check(self.lib.sphg_scene_get_sorted(self.h, ids.ctypes.data_as(C.POINTER(C.c_uint32)),
                                              cells.ctypes.data_as(C.POINTER(C.c_int32))))
        return ids, cells

    def neighbors(self, cap: int = None):
        cap = cap or (self.params.max_neighbors or 96)
        counts = np.zeros(self.n, dtype=np.uint32)
        ids = np.zeros((self.n, cap), dtype=np.uint32)
        _check(self.lib.sphg_scene_get_neighbors(self.h, counts.ctypes.data_as(C.POINTER(C.c_uint32)),
                                                 ids.ctypes.data_as(C.POINTER(C.c_uint32)), cap))
        return counts, ids

    def info(self):
        n, nb, nc = C.c_uint32(), C.c_uint32(), C.c_uint32()
        by = C.c_uint64()
        _check(self.lib.sphg_scene_info(self.h, C.byref(n), C.byref(nb), C.byref(nc), C.byref(by)))
        return {"n": n.value, "nb": nb.value, "ncells": nc.value, "device_bytes": by.value}
