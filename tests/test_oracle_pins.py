"""Pins of the FP64 oracle (no reference run exists: the reference ships no
tests or fixtures and cannot be compiled here).

Each check is chosen so that a plausible restatement mistake fails it:
closed forms on exact lattices (tests/golden/closed_forms.json), quadrature
normalisation, brute-force neighbor search, and invariants (antisymmetric
pressure pairs conserve momentum, a resting lattice stays at rest).
"""
import json
import math
import os

import numpy as np
import pytest

from sph_amd import configs, sphg
from sph_amd.sphg import SceneParams
import oracle.pyoracle as po

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "closed_forms.json")))


def _lattice(n, spacing, origin=0.0):
    g = np.arange(n) * spacing + origin
    return np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)


# --------------------------------------------------------------- kernels

@pytest.mark.parametrize("h", [0.1, 0.04, 0.02])
def test_cubic_W0_and_normalisation(h):
    assert math.isclose(po.kernel_W(h, 0.0), float(GOLD["cubic_W0_h3"]) / h ** 3, rel_tol=1e-14)
    # 4 pi int_0^h W(r) r^2 dr = 1 (composite Simpson on 4001 points)
    r = np.linspace(0.0, h, 4001)
    w = np.array([po.kernel_W(h, x) for x in r]) * r * r
    integral = 4 * math.pi * (h / 4000 / 3) * (w[0] + w[-1] + 4 * w[1:-1:2].sum() + 2 * w[2:-1:2].sum())
    assert abs(integral - 1.0) < 1e-10
    assert po.kernel_W(h, 1.0000001 * h) == 0.0


@pytest.mark.parametrize("q", [0.1, 0.3, 0.5, 0.6, 0.8, 0.99])
def test_cubic_gradient_is_dW_dr(q):
    h = 0.05
    r = q * h
    eps = 1e-7 * h
    dwdr = (po.kernel_W(h, r + eps) - po.kernel_W(h, r - eps)) / (2 * eps)
    g = po.kernel_gradW(h, [r, 0.0, 0.0])
    assert math.isclose(g[0], dwdr, rel_tol=1e-6)
    assert g[1] == 0.0 and g[2] == 0.0
    d = np.array([0.3, -0.5, 0.2])
    d = d / np.linalg.norm(d) * r
    g3 = po.kernel_gradW(h, d)
    assert np.allclose(g3, dwdr * d / r, rtol=1e-6)


def test_cubic_gradient_zero_at_origin_and_outside():
    h = 0.1
    assert np.all(po.kernel_gradW(h, [0.0, 0.0, 0.0]) == 0)
    assert np.all(po.kernel_gradW(h, [5e-6, 0.0, 0.0]) == 0)       # |r| <= 1e-5 cut
    assert np.all(po.kernel_gradW(h, [1.0001 * h, 0.0, 0.0]) == 0)


def test_spiky_and_poly6_gradients_closed_form():
    h = 0.04
    for r in (0.005, 0.02, 0.035):
        g = po.kernel_spiky_grad(h, [0.0, r, 0.0])
        assert math.isclose(g[1], -45.0 / (math.pi * h ** 6) * (h - r) ** 2, rel_tol=1e-13)
        p = po.kernel_poly6_grad(h, [0.0, 0.0, r])
        assert math.isclose(p[2], -945.0 / (32 * math.pi * h ** 9) * (h * h - r * r) ** 2 * r, rel_tol=1e-13)
    assert np.all(po.kernel_spiky_grad(h, [0.0, 0.0, 0.0]) == 0)


def test_cohesion_and_adhesion_as_written():
    h = 0.02
    c1 = 32.0 / (math.pi * h ** 9)
    # outer branch (h/2, h]
    r = 0.75 * h
    assert math.isclose(po.kernel_cohesion(h, r), c1 * (h - r) ** 3 * r ** 3, rel_tol=1e-13)
    # inner branch: the h^6/64 offset is NOT scaled by 32/(pi h^9)
    r = 0.25 * h
    assert math.isclose(po.kernel_cohesion(h, r), c1 * 2 * (h - r) ** 3 * r ** 3 - h ** 6 / 64, rel_tol=1e-13)
    # adhesion: (-4r^2/h + 6r - 2h)^(1/4) without its 0.007/h^3.25 factor
    r = 0.75 * h
    assert math.isclose(po.kernel_adhesion(h, r), (-4 * r * r / h + 6 * r - 2 * h) ** 0.25, rel_tol=1e-13)
    assert po.kernel_adhesion(h, 0.4 * h) == 0.0


# ------------------------------------------------------- lattice closed forms

def _params(method=sphg.DFSPH, r=0.025, **kw):
    p = SceneParams(method=method, particle_radius=r, gravity=(0.0, 0.0, 0.0))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def test_lattice_density_closed_form():
    r = 1.0 / 32.0                 # dyadic spacing: the FP32 lattice is exact
    x = _lattice(9, 2 * r, 1.0)
    far = np.array([[50.0, 50.0, 50.0]], dtype=np.float32)     # one boundary particle far away
    o = po.Oracle(_params(r=r), x, far)
    rho = o.get(po.OF_RHO)
    centre = np.argmin(np.linalg.norm(x - x.mean(0), axis=1))
    ratio = float(GOLD["lattice_density_ratio"])
    assert abs(rho[centre] / 1000.0 - ratio) < 1e-13
    # counts on the exact lattice
    cand = o.candidates(int(centre))
    d = np.linalg.norm(x[cand].astype(np.float64) - x[centre], axis=1) / (4 * r)
    assert int((d <= 1 + 1e-9).sum()) == GOLD["lattice_counts"]["candidates_q_le_1"]


def test_flat_wall_volume_closed_form():
    h = 0.125                      # dyadic: the sampled wall is an exact h/2 lattice
    b = sphg.sample_box((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), h)
    fluid = np.array([[0.5, 0.5, 0.5]], dtype=np.float32)
    o = po.Oracle(_params(r=h / 4), fluid, b)
    vb = o.get(po.OF_BVOL)
    centre = np.argmin(np.linalg.norm(b - np.array([0.5, 0.5, 0.0]), axis=1))
    assert abs(vb[centre] / h ** 3 - float(GOLD["wall_volume_over_h3"])) < 1e-13
    # a shell corner sees a quarter disk on each of its 3 faces:
    # sum W/k = 1 + 3 W(h/2)/k + 6 W(h/sqrt2)/k = 1.75 + 6 (1 - 1/sqrt2)^3
    corner = np.argmin(np.linalg.norm(b, axis=1))
    s_corner = 1.75 + 6 * (1 - math.sqrt(0.5)) ** 3
    assert abs(vb[corner] / h ** 3 - math.pi / (8 * s_corner)) < 1e-13


def test_pcisph_delta_closed_form():
    for r, dt in ((0.005, 2e-4), (0.025, 5e-4)):
        p = _params(method=sphg.PCISPH, r=r, time_step=dt)
        o = po.Oracle(p, np.array([[0.5, 0.5, 0.5]], np.float32), np.array([[9.0, 9.0, 9.0]], np.float32))
        h = 4 * r
        expect = float(GOLD["pcisph_delta_h2_dt2"]) * h * h / (dt * dt)
        assert math.isclose(o.scalars()["pcisph_delta"], expect, rel_tol=1e-12)


def test_single_particle_and_pair():
    r = 0.02
    h = 4 * r
    m = 1000.0 * (2 * r) ** 3
    far = np.array([[9.0, 9.0, 9.0]], np.float32)
    o = po.Oracle(_params(r=r), np.array([[0.5, 0.5, 0.5]], np.float32), far)
    assert math.isclose(o.get(po.OF_RHO)[0], m * 8 / (math.pi * h ** 3), rel_tol=1e-12)
    assert o.get(po.OF_ALPHA)[0] == 0.0     # no gradient: denominator 0 -> alpha 0
    pair = np.array([[0.5, 0.5, 0.5], [0.5 + h / 2, 0.5, 0.5]], np.float32)
    o2 = po.Oracle(_params(r=r), pair, far)
    k = 8 / (math.pi * h ** 3)
    hq = float(np.float32(0.5 + h / 2) - np.float32(0.5)) / h   # q of the FP32 pair
    w = 6 * hq ** 3 - 6 * hq ** 2 + 1 if hq <= 0.5 else 2 * (1 - hq) ** 3
    assert np.allclose(o2.get(po.OF_RHO), m * k * (1 + w), rtol=1e-12)


def test_tait_eos_closed_form():
    """WCSPH pressure p = max(0, kappa rho0/gamma ((rho/rho0)^gamma - 1))."""
    spec = configs.small_scene(sphg.WCSPH, n_side=6, jitter=0.03)
    o = po.Oracle(spec.params, spec.fluid, spec.boundary)
    o.step(1)
    rho = o.get(po.OF_RHO)
    p = o.get(po.OF_P)
    kap, gam, r0 = spec.params.stiffness, spec.params.gamma, spec.params.density0
    assert np.allclose(p, np.maximum(0.0, kap * r0 / gam * ((rho / r0) ** gam - 1)), rtol=1e-12, atol=1e-9)


# ----------------------------------------------------------------- grid

def test_grid_keys_buckets_candidates_bruteforce():
    spec = configs.small_scene(sphg.DFSPH, n_side=9, jitter=0.05)
    o = po.Oracle(spec.params, spec.fluid, spec.boundary)
    x = spec.fluid.astype(np.float64)
    h = spec.params.support_radius
    keys = np.floor(x / h).astype(np.int64)
    assert np.array_equal(o.keys(), keys)
    for i in range(0, spec.n, 7):
        pts = o.cell_points(i)
        exp = np.flatnonzero(np.all(keys == keys[i], axis=1))
        assert np.array_equal(pts, exp)                      # ascending id = insertion order
        cand = o.candidates(i)
        near = np.flatnonzero(np.all(np.abs(keys - keys[i]) <= 1, axis=1))
        assert sorted(cand.tolist()) == near.tolist()         # exactly the 27-cell union
        within = np.flatnonzero(np.linalg.norm(x - x[i], axis=1) <= h)
        assert set(within.tolist()) <= set(cand.tolist())    # no neighbor missed
        # concatenation order: x outermost, z innermost
        ck = keys[cand] - keys[i]
        code = (ck[:, 0] + 1) * 9 + (ck[:, 1] + 1) * 3 + (ck[:, 2] + 1)
        assert np.all(np.diff(code) >= 0)


# ----------------------------------------------------------- invariants

def test_pressure_pairs_conserve_momentum():
    """Fluid-fluid pressure terms are antisymmetric: sum_i m a_i = 0."""
    spec = configs.small_scene(sphg.WCSPH, n_side=7, jitter=0.05)
    far = np.array([[50.0, 50.0, 50.0]], np.float32)
    o = po.Oracle(spec.params, spec.fluid, far)
    o.step(1)
    a = o.get(po.OF_ACC_STEP)
    assert np.abs(a).max() > 1.0
    assert np.abs(a.sum(0)).max() < 1e-9 * np.abs(a).sum()


def test_resting_lattice_stays_at_rest():
    """No gravity, zero velocity, interior density below rho0: DFSPH must not
    move anything (rho* clamps to rho0, divergence error is zero)."""
    r = 0.025
    x = _lattice(8, 2 * r, 0.5)
    far = np.array([[50.0, 50.0, 50.0]], np.float32)
    o = po.Oracle(_params(r=r, time_step=1e-3, min_time_step=1e-6, max_time_step=5e-3, fps=25.0, cfl_factor=0.4,
                          eta=1e-4, eta_v=1e-3, min_iterations=2, max_iterations=100, min_iterations_v=1,
                          max_iterations_v=100), x, far)
    o.step(3)
    assert np.all(o.get(po.OF_V) == 0.0)
    assert np.array_equal(o.get(po.OF_X), x.astype(np.float32).astype(np.float64))
    assert np.all(o.get(po.OF_DIVERR) == 0.0)


def test_dfsph_min_iterations_and_dt_rule():
    spec = configs.small_scene(sphg.DFSPH, n_side=6)
    o = po.Oracle(spec.params, spec.fluid, spec.boundary)
    dt0 = spec.params.time_step
    o.step(1)
    s = o.scalars()
    assert s["iterations"] >= spec.params.min_iterations and s["iterations_v"] >= spec.params.min_iterations_v
    # first step: 0 previous iterations, cfl = 0.4*0.4*2r/max(0.1, sqrt|v|) > dt -> dt * 1.001
    assert s["cfl_time_step"] > dt0
    assert math.isclose(s["time_step"], dt0 * 1.001, rel_tol=1e-15)
    assert math.isclose(s["time"], s["time_step"], rel_tol=1e-15)


# ------------------------------------------- standard / artificial viscosity
# src/StandardViscosityForce.h:37-74, src/ArtificialViscosity.h:40-81, both
# restated in oracle/sph_oracle.cpp.  The acceleration of the step's
# non-pressure phase (OF_ACC_STEP of a DFSPH step, which has no other force)
# is checked against invariants a wrong sign, a missing (v_j - v_i)
# difference or a wrong vector would break.

def _visc_accel(method, vel, nu=0.05, nub=0.0, seed=3):
    spec = configs.small_scene(sphg.DFSPH, n_side=7, jitter=0.05, seed=seed, viscosity_method=method, viscosity=nu,
                               boundary_viscosity=nub)
    far = np.array([[50.0, 50.0, 50.0]], np.float32)
    o = po.Oracle(spec.params, spec.fluid, far, velocities=vel(spec.fluid.astype(np.float64)).astype(np.float32))
    o.step(1)
    a = o.get(po.OF_ACC_STEP)
    o.close()
    return spec, a


def _random_vel(x, seed=11):
    # multiples of 2^-10: exact in FP32, and so are their sums with the
    # dyadic shift below
    return np.round(np.random.default_rng(seed).uniform(-1.0, 1.0, size=x.shape) * 1024.0) / 1024.0


@pytest.mark.parametrize("method", [sphg.VISCOSITY_STANDARD, sphg.VISCOSITY_ARTIFICIAL])
def test_viscosity_galilean_invariant_and_linear(method):
    """The fluid terms depend on velocity differences only (a uniform shift
    changes nothing) and are linear in nu."""
    _, a = _visc_accel(method, _random_vel)
    _, a_shift = _visc_accel(method, lambda x: _random_vel(x) + np.array([0.25, -0.5, 0.125]))
    _, a2 = _visc_accel(method, _random_vel, nu=0.1)
    assert np.abs(a).max() > 1e-3
    assert np.abs(a_shift - a).max() < 1e-12 * np.abs(a).max()
    assert np.abs(a2 - 2.0 * a).max() < 1e-12 * np.abs(a).max() * 2


def test_artificial_viscosity_vanishes_for_rigid_rotation():
    """(v_i - v_j).r_ij = ((w x r_i) - (w x r_j)).r_ij = (w x r_ij).r_ij = 0;
    the FP32 rounding of w x r leaves ~1e-8 of the response to a random
    field (a formula mistake leaves O(1))."""
    w = np.array([0.3, -1.1, 0.7])
    _, a_rot = _visc_accel(sphg.VISCOSITY_ARTIFICIAL, lambda x: np.cross(w, x - 0.2))
    _, a_rnd = _visc_accel(sphg.VISCOSITY_ARTIFICIAL, _random_vel)
    assert np.abs(a_rot).max() < 1e-6 * np.abs(a_rnd).max()


@pytest.mark.parametrize("method", [sphg.VISCOSITY_STANDARD, sphg.VISCOSITY_ARTIFICIAL])
def test_viscosity_dissipates(method):
    """Both forms remove kinetic energy from a random velocity field:
    sum_i m a_i.v_i < 0 (pairwise -|v_i - v_j|^2 L_ij for the standard form,
    -(v_ij.r_ij)^2 |g_ij| for the artificial one, up to density factors)."""
    spec, a = _visc_accel(method, _random_vel)
    v = _random_vel(spec.fluid.astype(np.float64)).astype(np.float32).astype(np.float64)
    power = float(np.sum(a * v))
    assert power < 0.0
    assert abs(power) > 1e-3 * np.abs(a).sum() * np.abs(v).mean()


@pytest.mark.parametrize("method", [sphg.VISCOSITY_STANDARD, sphg.VISCOSITY_ARTIFICIAL])
def test_boundary_viscosity_opposes_motion(method):
    """With the walls of the box: the boundary term of a fluid moving as a
    whole opposes its velocity (the fluid terms vanish for uniform v)."""
    U = np.array([0.4, -0.2, 0.3])
    spec = configs.small_scene(sphg.DFSPH, n_side=7, jitter=0.0, viscosity_method=method, viscosity=0.05,
                               boundary_viscosity=0.05)
    vel = np.tile(U, (spec.n, 1)).astype(np.float32)
    o = po.Oracle(spec.params, spec.fluid, spec.boundary, velocities=vel)
    o.step(1)
    a = o.get(po.OF_ACC_STEP)
    o.close()
    near = np.abs(a).sum(1) > 0
    assert near.any() and not near.all()          # only particles near a wall feel it
    if method == sphg.VISCOSITY_STANDARD:
        # standard: -(nu_b/rho0) rho0 V_b L / rho_i v_i, exactly antiparallel to U
        cos = (a[near] @ U) / (np.linalg.norm(a[near], axis=1) * np.linalg.norm(U))
        assert np.all(cos < -1 + 1e-9)
    else:
        # artificial: sum_b c_b g_b (v.r)/(r^2 + eps) r with g_b < 0: a.v <= 0
        assert np.all(a[near] @ U <= 1e-12)
        assert (a[near] @ U).sum() < 0.0



# ------------------------------------------------------- cube boundary
# src/CubeBoundaryModel.cpp:16-94, restated in oracle cubeCorrect().

def test_cube_clamp_single_particle_closed_form():
    """One particle, no neighbors, no gravity, no pressure (its density m W(0)
    is rho0/pi < rho0): a WCSPH step moves it by v dt out of the box through
    the x and y minimum faces.  The clamp sets x, y to the face, and per axis
    in x, y, z order multiplies the outward component by -normalFct and the
    other two by tangentialFct, so the y branch rescales the vx the x branch
    already set: v = (-1, -2, 0.5) -> (0.5, -1.6, 0.4) -> (0.4, 0.8, 0.32)."""
    dt = 1e-3
    p = _params(r=0.025, method=sphg.WCSPH, gravity=(0.0, 0.0, 0.0), time_step=dt, min_time_step=dt,
                max_time_step=dt, boundary_method=sphg.BOUNDARY_CUBE, cube_enabled=1,
                cube_min=(0.4995, 0.499, 0.0), cube_max=(1.0, 1.0, 1.0), cube_normal_fct=0.5,
                cube_tangential_fct=0.8)
    x = np.array([[0.5, 0.5, 0.5]], np.float32)
    v = np.array([[-1.0, -2.0, 0.5]], np.float32)
    o = po.Oracle(p, x, np.zeros((0, 3), np.float32), velocities=v)
    o.step(1)
    xo, vo = o.get(po.OF_X)[0], o.get(po.OF_V)[0]
    assert xo[0] == 0.4995 and xo[1] == 0.499 and math.isclose(xo[2], 0.5 + 0.5 * dt, rel_tol=1e-15)
    assert np.allclose(vo, [0.4, 0.8, 0.32], rtol=1e-15, atol=0)
    # inside the box nothing changes
    o2 = po.Oracle(p, np.array([[0.7, 0.7, 0.5]], np.float32), np.zeros((0, 3), np.float32), velocities=v)
    o2.step(1)
    assert np.array_equal(o2.get(po.OF_V)[0], v[0].astype(np.float64))


def test_fp32_state_rounding_helpers():
    """The oracle's FP32 storage emulation (used to follow a free-running FP32
    solver): round_state_f32 leaves every position and velocity an FP32 value,
    a position on a cube face the nearest FP32 value inside the box, and keeps
    the densities; set_f32_positions makes DFSPH's updated positions FP32."""
    spec = configs.small_scene(sphg.DFSPH, n_side=6)
    f = spec.fluid.astype(np.float64)
    r = spec.params.particle_radius
    # faces a quarter ulp inside an FP32 value: the nearest FP32 lies outside
    zl = np.float32(f[:, 2].min() + 0.01 * r)
    xh = np.float32(f[:, 0].max() - 0.01 * r)
    lo = (f[:, 0].min() - 1.0, f[:, 1].min() - 1.0, float(zl) + 0.25 * float(np.spacing(zl)))
    hi = (float(xh) - 0.25 * float(np.spacing(xh)), f[:, 1].max() + 1.0, f[:, 2].max() + 1.0)
    assert np.float64(np.float32(hi[0])) > hi[0] and np.float64(np.float32(lo[2])) < lo[2]
    from dataclasses import replace
    params = replace(spec.params, boundary_method=sphg.BOUNDARY_CUBE, cube_enabled=1, cube_min=lo, cube_max=hi)
    empty = np.zeros((0, 3), np.float32)
    o = po.Oracle(params, spec.fluid, empty)
    o.step(1)
    x = o.get(po.OF_X)
    assert np.any(x[:, 0] == hi[0]) and np.any(x[:, 2] == lo[2])       # clamped onto the faces
    rho = o.get(po.OF_RHO)
    o.round_state_f32()
    x2, v2 = o.get(po.OF_X), o.get(po.OF_V)
    assert np.array_equal(x2, x2.astype(np.float32).astype(np.float64))
    assert np.array_equal(v2, v2.astype(np.float32).astype(np.float64))
    assert np.all(x2[:, 0] <= hi[0]) and np.all(x2[:, 2] >= lo[2])
    assert np.all(np.abs(x2 - x) <= np.spacing(np.abs(x).astype(np.float32)).astype(np.float64))
    assert np.array_equal(o.get(po.OF_RHO), rho)
    # without the cube, DFSPH's updated positions are FP32 values
    o2 = po.Oracle(spec.params, spec.fluid, spec.boundary)
    o2.set_f32_positions(True)
    o2.step(1)
    x3 = o2.get(po.OF_X)
    assert np.array_equal(x3, x3.astype(np.float32).astype(np.float64))
    o3 = po.Oracle(spec.params, spec.fluid, spec.boundary)
    o3.step(1)
    assert not np.array_equal(o3.get(po.OF_X), x3)
    assert np.abs(o3.get(po.OF_X) - x3).max() < 1e-6


# --------------------------------------------------------------- emitters

def _emitter_oracle(emitters, method=sphg.WCSPH, r=2.0 ** -7, dt=2.0 ** -10, gravity=(0.0, 0.0, 0.0)):
    p = _params(method=method, r=r, gravity=gravity, time_step=dt, min_time_step=dt, max_time_step=dt)
    empty = np.zeros((0, 3), np.float32)
    o = po.Oracle(p, empty, empty)
    sizes = [o.add_emitter(e) for e in emitters]
    return o, sizes


def test_emitter_square_group_exact():
    """Square group (src/Emitter.cpp:57-75): floor(w/dist) x floor(h/dist)
    at spacing dist = 2r, centred, x index outer and decreasing from +W/2.
    Emitted at t = start with zero offset and the identity rotation, the
    first group sits exactly at origin + group (dyadic values)."""
    r = 2.0 ** -7
    dist = 2 * r
    e = sphg.Emitter(type=sphg.EMITTER_SQUARE, num_particles=15, position=(0.5, 0.25, 0.125), velocity=0.75,
                     width=5 * dist, height=3 * dist)
    o, sizes = _emitter_oracle([e], r=r)
    assert sizes == [15]
    assert o.n == 0
    o.step(1)                        # WCSPH emits after the integration: no motion in the emitting step
    assert o.n == 15
    x, v = o.get(po.OF_X), o.get(po.OF_V)
    exp = np.array([[0.5 + 2 * dist - i * dist, 0.25 + dist - j * dist, 0.125] for i in range(5) for j in range(3)])
    assert np.array_equal(x, exp)
    assert np.array_equal(v, np.tile([0.0, 0.0, 0.75], (15, 1)))


def test_emitter_circle_group_closed_form():
    """Circle group (src/Emitter.cpp:31-55): the centre plus rings l = 1 ..
    floor((w - dist/2)/dist) of radius l dist holding floor(2 pi l) particles;
    the first particle of a ring is at angle 0, i.e. (0, radius)."""
    r = 2.0 ** -7
    dist = 2 * r
    e = sphg.Emitter(type=sphg.EMITTER_CIRCLE, num_particles=1000, width=3.4 * dist, velocity=1.0)
    o, sizes = _emitter_oracle([e], r=r)
    assert sizes == [1 + math.floor(2 * math.pi) + math.floor(4 * math.pi)] == [19]
    o.step(1)
    x = o.get(po.OF_X)
    rad = np.hypot(x[:, 0], x[:, 1])
    assert np.array_equal(x[0], [0.0, 0.0, 0.0])
    assert np.allclose(rad[1:7], dist, rtol=0, atol=1e-15) and np.allclose(rad[7:], 2 * dist, rtol=0, atol=1e-15)
    assert np.allclose(x[1], [0.0, dist, 0.0], atol=1e-18) and np.allclose(x[7], [0.0, 2 * dist, 0.0], atol=1e-18)
    assert np.all(x[:, 2] == 0.0)


def test_emitter_schedule_budget_and_time_offset():
    """Emission k is due at t_k = start + k spacing 1.105 2r / v
    (src/Emitter.cpp:85-97,131): it happens in the first step whose start
    time reaches t_k, one group per step, while a whole group is left in the
    budget.  Its positions carry the offset (v + g t) t, t = time - t_k."""
    r, dt = 2.0 ** -7, 2.0 ** -10
    dist = 2 * r
    interval = 10.37 * dt
    vel = 1.0 * 1.105 * dist / interval        # spacing 1
    g = (0.0, 0.0, -9.81)
    start = 3.3 * dt
    e = sphg.Emitter(type=sphg.EMITTER_SQUARE, num_particles=int(3.5 * 4), position=(0.0, 0.0, 0.0), velocity=vel,
                     width=2 * dist, height=2 * dist, start_time=start, spacing=1.0)
    o, sizes = _emitter_oracle([e], r=r, dt=dt, gravity=g)
    assert sizes == [4]
    due = [start + k * interval for k in range(3)]
    # step s starts at time s dt (fixed dt): emission k at the first s with s dt >= t_k
    exp_steps = [math.ceil(t / dt - 1e-12) for t in due]
    counts = []
    for s in range(60):
        o.step(1)
        counts.append(o.n)
    for k, s in enumerate(exp_steps):
        assert counts[s] == 4 * (k + 1), (k, s, counts[max(0, s - 1):s + 2])
        assert s == 0 or counts[s - 1] == 4 * k
    assert counts[-1] == 12                      # the half group left is never emitted
    # offset of the first group: emitted at step exp_steps[0], time exp_steps[0] dt
    o2, _ = _emitter_oracle([e], r=r, dt=dt, gravity=g)
    o2.step(exp_steps[0] + 1)
    t = exp_steps[0] * dt - start
    x = o2.get(po.OF_X)
    assert np.allclose(x[:, 2], (vel + g[2] * t) * t, rtol=1e-13, atol=0)
    assert np.allclose(np.sort(x[:, 0]), [-0.5 * dist, -0.5 * dist, 0.5 * dist, 0.5 * dist], rtol=0, atol=1e-15)


def test_emitter_rotation_quaternion():
    """rotation (w, x, y, z) rotates the group and the emission velocity:
    90 degrees about x maps local (x, y, 0) to (x, 0, y) and +z to -y."""
    r = 2.0 ** -7
    dist = 2 * r
    c = math.cos(math.pi / 4)
    e = sphg.Emitter(type=sphg.EMITTER_SQUARE, num_particles=6, position=(1.0, 2.0, 3.0), velocity=0.5,
                     rotation=(c, c, 0.0, 0.0), width=3 * dist, height=2 * dist)
    o, _ = _emitter_oracle([e], r=r)
    o.step(1)
    x, v = o.get(po.OF_X), o.get(po.OF_V)
    loc = np.array([[dist - i * dist, 0.5 * dist - j * dist] for i in range(3) for j in range(2)])
    exp = np.stack([1.0 + loc[:, 0], 2.0 + 0.0 * loc[:, 1], 3.0 + loc[:, 1]], -1)
    assert np.allclose(x, exp, rtol=0, atol=1e-15)
    assert np.allclose(v, np.tile([0.0, -0.5, 0.0], (6, 1)), rtol=0, atol=1e-15)


# ------------------------------------------------- PCISPH boundary method

def _cubic_np(r, h):
    q = r / h
    k = 8.0 / (math.pi * h ** 3)
    return np.where(q <= 0.5, k * (6 * q ** 3 - 6 * q ** 2 + 1), np.where(q <= 1.0, k * 2 * (1 - q) ** 3, 0.0))


def _pb_scene(method, fluid, vel, dt=2.0 ** -6, **kw):
    r = 1.0 / 32.0
    h = 4 * r
    pts, nrm = sphg.sample_box_normals((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), h, True)
    p = _params(method=method, r=r, time_step=dt, min_time_step=dt, max_time_step=dt, boundary_method=1, **kw)
    o = po.Oracle(p, np.asarray(fluid, np.float32), pts, velocities=np.asarray(vel, np.float32),
                  boundary_normals=nrm)
    return o, pts.astype(np.float64), nrm.astype(np.float64), h


def test_pcisph_box_sampler_normals():
    """PCISPHBoundaryModel::sampleCube/addCube (src/PCISPHBoundaryModel.cpp:
    7-27, 60-189): the product sampler equals the oracle's restatement bit
    for bit (as FP32); spacing ceil(l / 0.45h); unit normals pointing into
    an inverted box (faces, edges (1,1,0)/sqrt2, corners (1,1,1)/sqrt3)."""
    for mn, mx, h, inv in (((0, 0, 0), (1, 1, 1), 0.125, True), ((0.1, -0.2, 0.3), (0.9, 0.45, 1.7), 0.1, False)):
        p, n = sphg.sample_box_normals(mn, mx, h, inv)
        po_p, po_n = po.sample_box_normals(mn, mx, h, inv)
        assert np.array_equal(p, po_p.astype(np.float32)) and np.array_equal(n, po_n.astype(np.float32))
        cnt = [math.ceil(abs(b - a) / (0.45 * h)) for a, b in zip(mn, mx)]
        tot = (cnt[0] + 1) * (cnt[1] + 1) * (cnt[2] + 1) - (cnt[0] - 1) * (cnt[1] - 1) * (cnt[2] - 1)
        assert len(p) == tot
        assert np.allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-7)
        c = 0.5 * (np.asarray(mn) + np.asarray(mx))
        inward = np.einsum("ij,ij->i", n.astype(np.float64), c - p) > 0
        assert np.all(inward == inv)
        faces = np.sum(np.isclose(np.abs(n), 1.0, atol=1e-7), axis=1) == 1
        corners = np.all(np.isclose(np.abs(n), 1 / math.sqrt(3), atol=1e-7), axis=1)
        assert faces.sum() > 0 and corners.sum() == 8
    # sphere: the Akinci sampler's points, normals inv (center - point)
    ps, ns = sphg.sample_sphere_normals((0.3, 0.4, 0.5), 0.2, 0.02, False)
    assert np.array_equal(ps, sphg.sample_sphere((0.3, 0.4, 0.5), 0.2, 0.02))
    assert np.allclose(ns, po.sphere_normals(ps.astype(np.float64), (0.3, 0.4, 0.5), False), atol=1e-7)


def test_pcisph_boundary_projection_flat_wall():
    """A particle moving into the x = 1 face (inverted box, no forces): after
    the integration it lies within h/2 of the face particles and is moved
    along the face normal (-1, 0, 0) by sum w (h/2 - |q|) / sum w, w =
    (h/2 - |q|) / (h/2) (src/PCISPHBoundaryModel.cpp:282-307); the velocity
    becomes tangentialFct v_t - normalFct v_n (:339-355)."""
    x0 = np.array([[0.965, 0.5, 0.5]])
    v0 = np.array([[1.0, 0.3, 0.0]])
    o, pts, nrm, h = _pb_scene(sphg.WCSPH, x0, v0, stiffness=0.0, pb_normal_fct=0.5, pb_tangential_fct=0.8)
    dt = 2.0 ** -6
    o.step(1)
    v32 = v0.astype(np.float32).astype(np.float64)[0]
    x1 = x0.astype(np.float32).astype(np.float64)[0] + v32 * dt
    R = 0.5 * h
    q = np.linalg.norm(x1 - pts, axis=1)
    hit = q < R
    assert hit.sum() >= 4 and np.allclose(nrm[hit], [-1.0, 0.0, 0.0])
    w = (R - q[hit]) / R
    s = (w * (R - q[hit])).sum() / w.sum()
    assert np.allclose(o.get(po.OF_X)[0], x1 + s * np.array([-1.0, 0.0, 0.0]), rtol=0, atol=1e-14)
    assert np.allclose(o.get(po.OF_V)[0], [-0.5 * v32[0], 0.8 * v32[1], 0.0], rtol=0, atol=1e-15)
    assert np.allclose(o.get(po.OF_PBN)[0], [-1.0, 0.0, 0.0], atol=1e-15)


def test_pcisph_boundary_needs_an_existing_cell():
    """The projection looks the position's cell up in the hash table: a cell
    no particle was ever inserted in has no boundary candidates, so a
    particle that enters one within h/2 of the wall is not projected in that
    step (and its normal is NaN); once inserted there it is.  The same
    motion starting inside that cell is projected at once."""
    dt = 2.0 ** -6
    # from cell 6 (x < 0.875) into cell 7 (0.875 <= x < 1), 0.056 from the wall
    o, pts, _, h = _pb_scene(sphg.WCSPH, [[0.874, 0.5, 0.5]], [[4.5, 0.0, 0.0]], stiffness=0.0)
    o.step(1)
    x1 = np.float64(np.float32(0.874)) + 4.5 * dt
    assert x1 >= 0.875 and 1.0 - x1 < 0.5 * h
    assert o.get(po.OF_X)[0, 0] == x1 and np.all(np.isnan(o.get(po.OF_PBN)[0]))
    o.step(1)                     # inserted in cell 7 last step: now projected
    assert o.get(po.OF_X)[0, 0] < x1 + 4.5 * dt and np.allclose(o.get(po.OF_PBN)[0], [-1, 0, 0])
    # control: starts in cell 7 (visited at the first insert)
    o2, _, _, _ = _pb_scene(sphg.WCSPH, [[0.876, 0.5, 0.5]], [[4.5, 0.0, 0.0]], stiffness=0.0)
    o2.step(1)
    assert o2.get(po.OF_X)[0, 0] < np.float64(np.float32(0.876)) + 4.5 * dt
    assert np.allclose(o2.get(po.OF_PBN)[0], [-1, 0, 0])


@pytest.mark.parametrize("method", [sphg.WCSPH, sphg.PCISPH])
def test_pcisph_boundary_densities_brute_force(method):
    """Boundary densities of the PCISPH method with the fluid far from the
    wall: WCSPH sums m_b W (cubic) over the boundary (src/SPHSolver.cpp:
    93-121) and its pressure is the Tait EOS; the PCISPH loop sums m_b Poly6
    (src/PCISPHSolver.cpp:214-247).  Brute force over all pairs."""
    o, pts, _, h = _pb_scene(method, [[0.5, 0.5, 0.5]], [[0.0, 0.0, 0.0]], stiffness=1000.0, gamma=7.0,
                             eta=0.01, min_iterations=2)
    o.step(1)
    bd = o.get(po.OF_BDENS)
    far = np.linalg.norm(pts - 0.5, axis=1) > 1.5 * h
    mb = 1000.0 * (2 * (1 / 32)) ** 3
    sel = np.flatnonzero(far)[::37]
    d = np.linalg.norm(pts[sel, None, :] - pts[None, :, :], axis=2)
    if method == sphg.WCSPH:
        exp = (mb * _cubic_np(d, h)).sum(1)
        assert np.allclose(bd[sel], exp, rtol=1e-12)
        bp = o.get(po.OF_BPRESS)
        assert np.allclose(bp[sel], np.maximum(0.0, 1000.0 * 1000.0 / 7.0 * ((exp / 1000.0) ** 7 - 1)), rtol=1e-10)
    else:
        c6 = 315.0 / (64 * math.pi * h ** 9)
        exp = (mb * np.where(d <= h, c6 * (h * h - d * d) ** 3, 0.0)).sum(1)
        assert np.allclose(bd[sel], exp, rtol=1e-12)


def test_pcisph_boundary_normals_nan_without_collision():
    o, _, _, _ = _pb_scene(sphg.DFSPH, [[0.5, 0.5, 0.5], [0.52, 0.5, 0.5]], [[0, 0, 0], [0, 0, 0]])
    o.step(1)
    assert np.all(np.isnan(o.get(po.OF_PBN)))


# ------------------------------------------------ whole-step invariants
# Full solver steps -- every phase, both DFSPH loops, the PCISPH loop, the
# Akinci surface tension -- against what the physics fixes for any correct
# implementation: a mirror-symmetric state stays mirror-symmetric (a swapped
# or transposed vector component breaks it), free fluid conserves momentum
# (a pressure or cohesion term that is not antisymmetric in i, j breaks it),
# and a uniform velocity added to every particle only translates the result
# (an absolute velocity where a relative one belongs breaks it).  The states
# are dyadic (exact in FP32), and the one boundary particle is far away.
# PCISPH's predicted densities take their candidates from the cells of the
# current positions (src/PCISPHSolver.cpp:181), so they depend on where the
# particles sit in the fixed grid: the mirror plane lies on a cell boundary
# and the velocity shift moves every particle by exactly one cell per step.

_FAR = np.array([[50.0, 50.0, 50.0]], np.float32)
_Q = 2.0 ** -20


def _sym_block(n=8, r=1.0 / 64.0, seed=21):
    """A jittered lattice block mirror-symmetric about the plane x = c = 10 h
    (a cell boundary, h = 4r), with mirrored velocities (vx -> -vx):
    particle k and mirror[k] are images.  No particle is within 2r/2 of a
    cell boundary."""
    s = 2.0 * r
    rng = np.random.default_rng(seed)
    c = 10.0 * 4.0 * r
    g = c - (n - 1) * s / 2.0 + np.arange(n) * s
    xs, vs, idx = [], [], {}
    for ix in range(n // 2):
        for iy in range(n):
            for iz in range(n):
                j = np.round(rng.uniform(-0.05, 0.05, 3) * s / _Q) * _Q
                v = np.round(rng.uniform(-1.0, 1.0, 3) * 1024.0) / 1024.0
                p = np.array([g[ix], g[iy], g[iz]]) + j
                idx[(ix, iy, iz)] = len(xs)
                xs.append(p)
                vs.append(v)
                idx[(n - 1 - ix, iy, iz)] = len(xs)
                xs.append(np.array([2.0 * c - p[0], p[1], p[2]]))
                vs.append(np.array([-v[0], v[1], v[2]]))
    x, v = np.array(xs), np.array(vs)
    assert np.array_equal(x.astype(np.float32).astype(np.float64), x)
    mirror = np.arange(len(x)) ^ 1
    return x, v, mirror, c


def _solver_params(method, r, xsph=True, **kw):
    """DFSPH; WCSPH with XSPH viscosity (unless xsph is False: its m/rho_j
    weights are not antisymmetric, so it does not conserve momentum); PCISPH
    with Akinci surface tension."""
    base = dict(time_step=5e-4, min_time_step=1e-6, max_time_step=5e-3, cfl_factor=0.4, eta=1e-4, eta_v=1e-3,
                min_iterations=2, max_iterations=100, min_iterations_v=1, max_iterations_v=100)
    if method == sphg.WCSPH and xsph:
        base.update(viscosity_method=sphg.VISCOSITY_XSPH, viscosity=0.01)
    if method == sphg.PCISPH:
        base.update(eta=0.01, min_iterations=3, surface_tension_method=sphg.SURFACE_TENSION_AKINCI,
                    surface_tension=0.05)
    base.update(kw)
    return _params(method=method, r=r, **base)


_METHODS = [sphg.DFSPH, sphg.WCSPH, sphg.PCISPH]


@pytest.mark.parametrize("method", _METHODS)
def test_whole_steps_keep_mirror_symmetry(method):
    x, v, mirror, c = _sym_block()
    o = po.Oracle(_solver_params(method, 1.0 / 64.0, gravity=(0.0, 0.0, -9.81)), x, _FAR, velocities=v)
    o.step(3)
    xa, va = o.get(po.OF_X), o.get(po.OF_V)
    assert np.abs(va - v).max() > 1e-2          # the solver moved things
    img_x = np.column_stack([2.0 * c - xa[:, 0], xa[:, 1], xa[:, 2]])
    img_v = np.column_stack([-va[:, 0], va[:, 1], va[:, 2]])
    assert np.abs(xa[mirror] - img_x).max() < 1e-12
    assert np.abs(va[mirror] - img_v).max() < 1e-9 * np.abs(va).max()
    o.close()


@pytest.mark.parametrize("method", _METHODS)
def test_whole_steps_conserve_momentum(method):
    x, v, _, _ = _sym_block(seed=22)
    v = v * np.array([1.0, 0.5, 0.25])           # break the mirror: net momentum is nonzero
    o = po.Oracle(_solver_params(method, 1.0 / 64.0, xsph=False), x, _FAR, velocities=v)
    p0 = v.sum(0)
    o.step(3)
    va = o.get(po.OF_V)
    assert np.abs(va - v).max() > 1e-2
    assert np.abs(va.sum(0) - p0).max() < 1e-11 * np.abs(va).sum()
    o.close()


@pytest.mark.parametrize("method", _METHODS)
def test_whole_steps_galilean_invariant(method):
    x, v, _, _ = _sym_block(seed=23)
    dt = 2.0 ** -11
    u0 = np.array([(1.0 / 16.0) / dt, 0.0, 0.0])     # one cell (h = 4r = 1/16) per step
    p = _solver_params(method, 1.0 / 64.0, gravity=(0.0, 0.0, -9.81), time_step=dt, min_time_step=dt,
                       max_time_step=dt)
    a = po.Oracle(p, x, _FAR, velocities=v)
    b = po.Oracle(p, x, _FAR, velocities=v + u0)
    a.step(3)
    b.step(3)
    t = a.scalars()["time"]
    assert t == 3 * dt and t == b.scalars()["time"]
    assert a.scalars()["iterations"] == b.scalars()["iterations"]
    va, vb = a.get(po.OF_V), b.get(po.OF_V)
    assert np.abs(va - v).max() > 1e-2
    assert np.abs(vb - u0 - va).max() < 1e-9 * np.abs(va).max()
    assert np.abs(b.get(po.OF_X) - u0 * t - a.get(po.OF_X)).max() < 1e-12
    a.close()
    b.close()


def test_divergence_error_20_neighbor_cutoff():
    """computeDivergenceError zeroes the error of a particle with fewer than
    20 fluid neighbors of nonzero kernel gradient (src/DFSPHSolver.cpp:446-447,
    480-483).  Both loops are skipped (min 0 iterations, huge tolerances)
    and k starts at 0, so the step only moves x by v dt and the error is
    formed once from the new positions.  A converging velocity field makes
    every error positive before the cutoff; the neighbor counts come from a
    brute-force FP64 distance test on the returned positions."""
    r = 1.0 / 64.0
    s = 2.0 * r
    h = 4.0 * r
    rng = np.random.default_rng(31)
    x = _lattice(6, s, 0.5) + np.round(rng.uniform(-0.15, 0.15, (216, 3)) * s / _Q) * _Q
    centre = x.mean(0)
    v = np.round(-(x - centre) * 0.5 / _Q) * _Q
    p = _params(r=r, time_step=5e-4, min_time_step=5e-4, max_time_step=5e-4, cfl_factor=0.4, eta=1e9, eta_v=1e9,
                min_iterations=0, max_iterations=100, min_iterations_v=0, max_iterations_v=100)
    o = po.Oracle(p, x, _FAR, velocities=v)
    o.step(1)
    assert o.scalars()["iterations"] == 0 and o.scalars()["iterations_v"] == 0
    xa = o.get(po.OF_X)
    de = o.get(po.OF_DIVERR)
    d = np.linalg.norm(xa[:, None, :] - xa[None, :, :], axis=2)
    q = d / h
    cnt = ((d > 1e-5) & (q < 1.0)).sum(1)
    assert set(range(17, 24)) <= set(cnt.tolist())        # counts on both sides of the cutoff
    assert np.all(de[cnt < 20] == 0.0)
    assert np.all(de[cnt >= 20] > 0.0)
    o.close()


def test_warm_start_eps_branch_linearity_and_clamp():
    """densityWarmStart (src/DFSPHSolver.cpp:602-689): k_i <- min(k_i/rho_i,
    1e-9)/dt^2, then v_i -= dt sum_j m (k_i + k_j) gradW_ij, a term kept only
    where |k_i + k_j| > 1e-5.  With both loops skipped and v = 0 it is the
    step's only velocity change, so: below the eps branch nothing moves;
    between eps and the clamp the change is linear in k (doubling k doubles
    it bit for bit); above the clamp every k gives the same change, which is
    the change of k_i just below 1e-9 rho_i to within its 1% offset."""
    r = 1.0 / 64.0
    s = 2.0 * r
    rng = np.random.default_rng(41)
    x = _lattice(6, s, 0.5) + np.round(rng.uniform(-0.15, 0.15, (216, 3)) * s / _Q) * _Q
    v0 = np.zeros_like(x)
    dt = 5e-4
    p = _params(r=r, time_step=dt, min_time_step=dt, max_time_step=dt, cfl_factor=0.4, eta=1e9, eta_v=1e9,
                min_iterations=0, max_iterations=100, min_iterations_v=0, max_iterations_v=100)
    o = po.Oracle(p, x, _FAR)
    rho = o.get(po.OF_RHO)
    assert rho.min() > 400.0 and rho.max() < 1200.0    # keeps k = 1e-10 below and 1e-8 above the eps branch

    def dv(k):
        o.set_state(x, v0, k=np.broadcast_to(np.asarray(k, np.float32), (len(x),)).copy(), time_step=dt)
        o.step(1)
        assert o.scalars()["iterations"] == 0 and o.scalars()["iterations_v"] == 0
        return o.get(po.OF_V)

    assert np.all(dv(1e-10) == 0.0)                       # 2 k/rho/dt^2 < 1e-5: every term dropped
    a1, a2 = dv(1e-8), dv(2e-8)
    assert np.abs(a1).max() > 1e-9
    assert np.array_equal(a2, 2.0 * a1)
    c1, c2 = dv(1e-3), dv(1.0)
    assert np.array_equal(c1, c2) and np.abs(c1).max() > 30.0 * np.abs(a1).max()   # 1e-9 against ~1e-11
    below = dv((0.99e-9 * rho).astype(np.float32))
    rel = np.abs(c1 - below).max() / np.abs(c1).max()
    assert 0.005 < rel < 0.02
    o.close()


def _cubic_grad_np(d, h):
    """Cubic spline gradient (src/CubicSpline.h) for difference vectors d
    (..., 3): 48/(pi h^3) q (3q - 2) r/(|r| h) for q <= 1/2, -48/(pi h^3)
    (1 - q)^2 r/(|r| h) for q <= 1, 0 for |r| <= 1e-5 and q > 1."""
    rl = np.linalg.norm(d, axis=-1)
    q = rl / h
    l = 48.0 / (math.pi * h ** 3)
    f = np.where(q <= 0.5, l * q * (3.0 * q - 2.0), -l * (1.0 - q) ** 2)
    f = np.where((rl > 1e-5) & (q <= 1.0), f, 0.0)
    return (f / np.where(rl > 0, rl * h, 1.0))[..., None] * d


def test_divergence_error_of_given_velocities_brute_force():
    """oracle_divergence_error: computeDivergenceError (src/DFSPHSolver.cpp:
    408-492) with given velocities, against a brute-force sum over all fluid
    and boundary pairs: sum_j m (v_i - v_j).gradW_ij + sum_b rho0 V_b
    v_i.gradW_ib, zero below 20 fluid neighbors of nonzero gradient, else
    clamped at 0.  The oracle's own velocities are kept."""
    spec = configs.small_scene(sphg.DFSPH, n_side=8)
    p = spec.params
    o = po.Oracle(p, spec.fluid, spec.boundary)
    o.step(2)
    v_own = o.get(po.OF_V)
    x = o.get(po.OF_X)
    xb = np.asarray(spec.boundary, dtype=np.float64)
    vb = o.get(po.OF_BVOL)
    rng = np.random.default_rng(17)
    vel = (rng.standard_normal((spec.n, 3)) * 0.3).astype(np.float32)
    got = o.divergence_error(vel)
    assert np.array_equal(o.get(po.OF_V), v_own)
    h = p.support_radius
    m = p.density0 * (2.0 * p.particle_radius) ** 3
    v = vel.astype(np.float64)
    g = _cubic_grad_np(x[:, None, :] - x[None, :, :], h)                       # (n, n, 3)
    de = m * np.einsum("ijk,ijk->i", v[:, None, :] - v[None, :, :], g)
    gb = _cubic_grad_np(x[:, None, :] - xb[None, :, :], h)                     # (n, nb, 3)
    de += p.density0 * np.einsum("ijk,ik->i", (vb[None, :, None] * gb), v)
    cnt = (np.linalg.norm(g, axis=-1) > 0).sum(1)
    exp = np.where(cnt < 20, 0.0, np.maximum(de, 0.0))
    assert (cnt < 20).any() and (cnt >= 20).any() and (exp > 0).any()
    assert np.abs(got - exp).max() <= 1e-10 * np.abs(exp).max()
    o.close()
