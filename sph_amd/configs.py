"""Scene builders for the benchmark workloads and the tests.

Geometry resolutions (the workloads as written do not fit their boxes):
  * WCSPH dam break, 32^3 at spacing 0.05: box 3 x 1.7 x 1.7 (a 32-particle
    edge is 1.6 m), first lattice layer 2.5 r from the walls.
  * DFSPH dam break, 100^3 at spacing 0.02: box 4 x 2.03 x 2.03, first layer
    2.5 r from every wall, single-layer Akinci walls (102,410 particles).
    A layer 2 r from an Akinci wall starts 15 % over-compressed at the
    corners (q = 1/2 to the wall), which pins the first DFSPH steps at the
    100-iteration cap; 2.5 r starts at or below rest density.
  * PCISPH spheres: lattice sites of a 2 x 2 x 2 block lattice inside any of
    64 seeded spheres, container box and slab in one boundary model.
  * Batched DFSPH: spacing 0.02, box (2L, L, L) of volume N s^3 / fill.
Gravity is z-up (0, 0, -9.81), the Blender add-on's convention.  Seeded
jitter is +-1% of the spacing per coordinate (SplitMix64).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, replace
from typing import List, Tuple

import numpy as np

from . import sphg
from .sphg import DFSPH, PCISPH, WCSPH, SceneParams

G = (0.0, 0.0, -9.81)


@dataclass
class SceneSpec:
    name: str
    params: SceneParams
    fluid: np.ndarray      # (n, 3) float32
    boundary: np.ndarray   # (nb, 3) float32

    @property
    def n(self) -> int:
        return int(self.fluid.shape[0])

    @property
    def nb(self) -> int:
        return int(self.boundary.shape[0])


class SplitMix64:
    """SplitMix64 stream; u = (x >> 11) * 2^-53."""

    def __init__(self, seed: int):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next(self) -> int:
        self.s = (self.s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next() >> 11) * (1.0 / 9007199254740992.0)

    def log_uniform_int(self, a: float, b: float) -> int:
        return int(round(math.exp(math.log(a) + self.uniform() * (math.log(b) - math.log(a)))))


def dam_break(method: int, n_side, spacing: float, box, jitter: float = 0.01, seed: int = 1, gap: float = 2.0,
              **overrides) -> SceneSpec:
    """Lattice block of n_side (int or 3-tuple) particles per edge in the box
    corner, Akinci box walls; the first lattice layer sits `gap` radii from
    the walls (2 = block corner at (r, r, r))."""
    r = spacing / 2.0
    ns = (n_side,) * 3 if isinstance(n_side, int) else tuple(n_side)
    mn = ((gap - 1.0) * r,) * 3
    mx = tuple(mn[a] + ns[a] * spacing for a in range(3))
    fluid = sphg.fluid_block(mn, mx, r, jitter, seed)
    h = 4.0 * r
    boundary = sphg.sample_box((0.0, 0.0, 0.0), box, h)
    p = SceneParams(method=method, particle_radius=r, gravity=G)
    p = replace(p, **overrides)
    return SceneSpec(f"dam_{method}_{len(fluid)}", p, fluid, boundary)


def wcsph_dam_32k(**kw) -> SceneSpec:
    """WCSPH dam break, 32^3 = 32,768 particles, spacing 0.05, stiffness 1000,
    gamma 7, XSPH 0.01, fixed dt 5e-4."""
    s = dam_break(WCSPH, 32, 0.05, (3.0, 1.7, 1.7), seed=1, gap=2.5, stiffness=1000.0, gamma=7.0,
                  viscosity_method=sphg.VISCOSITY_XSPH, viscosity=0.01, boundary_viscosity=0.0,
                  time_step=5e-4, min_time_step=5e-4, max_time_step=5e-4, **kw)
    s.name = "wcsph_dam_32k"
    return s


def dfsph_dam_1m(n_side: int = 100, **kw) -> SceneSpec:
    """DFSPH dam break, 100^3 = 1,000,000 particles, spacing 0.02, box
    4 x 2.03 x 2.03, CFL 0.4, max dt 5e-3, eta 0.01 %, eta_V 0.1 %, max 100
    iterations."""
    side = n_side * 0.02 + 0.03
    box = (4.0 * n_side / 100.0, side, side)
    s = dam_break(DFSPH, n_side, 0.02, box, seed=2, gap=2.5, cfl_factor=0.4, time_step=1e-3, max_time_step=5e-3,
                  min_time_step=1e-6, eta=1e-4, eta_v=1e-3, min_iterations=2, max_iterations=100,
                  min_iterations_v=1, max_iterations_v=100, **kw)
    s.name = f"dfsph_dam_{s.n}"
    return s


def pcisph_spheres(nspheres: int = 64, spacing: float = 0.01, box: float = 2.0, seed: int = 3,
                   rmin: float = 0.05, rmax: float = 0.15, **kw) -> SceneSpec:
    """PCISPH with Akinci surface tension (0.05) and adhesion (0.1): fluid
    spheres dropped in a box over a 0.5 x 0.5 x 0.1 slab; eta 1 %, 3 minimum
    iterations, dt 2e-4."""
    rng = SplitMix64(seed)
    centers, radii = [], []
    for _ in range(nspheres):
        c = [rng.uniform() * box for _ in range(3)]
        centers.append(c)
        radii.append(rmin + rng.uniform() * (rmax - rmin))
    r = spacing / 2.0
    h = 4.0 * r
    fluid = sphg.fluid_spheres((0.0, 0.0, 0.0), (box, box, box), r, centers, radii)
    s = box / 2.0
    slab_min = np.array([s - 0.25 * box / 2.0, s - 0.25 * box / 2.0, 0.2 * box])
    slab_max = slab_min + np.array([0.25 * box, 0.25 * box, 0.05 * box])
    # keep the fluid 3 radii clear of the slab and the container walls (an
    # Akinci wall closer than that over-compresses the first layer)
    c = 3.0 * r
    inside = np.all((fluid >= slab_min - c) & (fluid <= slab_max + c), axis=1)
    near_wall = np.any((fluid < c) | (fluid > box - c), axis=1)
    fluid = np.ascontiguousarray(fluid[~inside & ~near_wall])
    walls = sphg.sample_box((0.0, 0.0, 0.0), (box, box, box), h)
    slab = sphg.sample_box(tuple(slab_min), tuple(slab_max), h)
    boundary = np.ascontiguousarray(np.concatenate([walls, slab]))
    p = SceneParams(method=PCISPH, particle_radius=r, gravity=G, time_step=2e-4, min_time_step=2e-4,
                    max_time_step=2e-4, eta=0.01, min_iterations=3, max_iterations=100,
                    surface_tension_method=sphg.SURFACE_TENSION_AKINCI, surface_tension=0.05,
                    adhesion_method=sphg.ADHESION_AKINCI, adhesion=0.1)
    p = replace(p, **kw)
    return SceneSpec(f"pcisph_spheres_{len(fluid)}", p, fluid, boundary)


def batched_dfsph(nscenes: int = 256, nmin: int = 4096, nmax: int = 262144, seed: int = 4,
                  spacing: float = 0.02) -> List[SceneSpec]:
    """Independent DFSPH scenes, N log-uniform in [nmin, nmax], each a jittered
    lattice block filling U[0.2, 0.6] of a (2L, L, L) box."""
    rng = SplitMix64(seed)
    out = []
    for s in range(nscenes):
        n_target = rng.log_uniform_int(nmin, nmax)
        fill = 0.2 + 0.4 * rng.uniform()
        L = (n_target * spacing ** 3 / (2.0 * fill)) ** (1.0 / 3.0)
        f = fill ** (1.0 / 3.0)
        # block edge counts with the box aspect, product close to n_target
        ny = max(2, int(round((n_target / 2.0) ** (1.0 / 3.0))))
        nx = max(2, int(round(n_target / (ny * ny))))
        box = (max(2 * L, (nx + 1) * spacing + 0.02), max(L, (ny + 1) * spacing + 0.02),
               max(L, (ny + 1) * spacing + 0.02 * f))
        spec = dam_break(DFSPH, (nx, ny, ny), spacing, box, seed=1000 + s, gap=2.5, cfl_factor=0.4, time_step=1e-3,
                         max_time_step=5e-3, min_time_step=1e-6, eta=1e-4, eta_v=1e-3, min_iterations=2,
                         max_iterations=100, min_iterations_v=1, max_iterations_v=100)
        spec.name = f"batch{s}_{spec.n}"
        out.append(spec)
    return out


def small_scene(method: int, n_side: int = 8, spacing: float = 0.05, jitter: float = 0.01, seed: int = 7,
                box=None, gap: float = 2.5, **kw) -> SceneSpec:
    """A small dam break for parity tests (fluid block in a box corner, first
    layer 2.5 radii from the walls so the start is not over-compressed)."""
    if box is None:
        L = n_side * spacing
        box = (2.5 * L + 2 * spacing, L + gap * spacing, 1.5 * L + 2 * spacing)
    base = dict(time_step=1e-3 if method == DFSPH else 5e-4, max_time_step=5e-3, min_time_step=1e-6)
    if method == WCSPH:
        base.update(stiffness=1000.0, gamma=7.0, min_time_step=5e-4, max_time_step=5e-4)
    if method == PCISPH:
        base.update(eta=0.01, min_iterations=3, time_step=5e-4, min_time_step=5e-4, max_time_step=5e-4)
    if method == DFSPH:
        base.update(cfl_factor=0.4, eta=1e-4, eta_v=1e-3, min_iterations=2, max_iterations=100,
                    min_iterations_v=1, max_iterations_v=100)
    base.update(kw)
    return dam_break(method, n_side, spacing, box, jitter=jitter, seed=seed, gap=gap, **base)
