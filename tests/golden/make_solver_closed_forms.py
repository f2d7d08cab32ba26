"""Known-answer cases for the solver magnitudes (no reference run needed).

Writes tests/golden/solver_closed_forms.json.  Each case is a scene of 1 to
33 fluid particles at dyadic (FP32-exact) positions, small enough that one
solver step can be written out by hand: the expected fields below follow
from the reference's formulas evaluated in 40-digit decimal arithmetic,
term by term, with none of the oracle's code.  The scenes are chosen so
that the mathematics fixes the answer independently of summation order
(symmetric sums vanish, single pairs have one term) or reduces it to a
handful of explicit terms:

  dfsph_pair_q*       two approaching particles, one density iteration:
                      alpha (src/DFSPHSolver.cpp:171-186), rho* (:272-297),
                      kappa = (rho* - rho0) alpha / (dt^2 rho) (:346),
                      k += (rho* - rho0) alpha (:349), the velocity sweep
                      (:353-394), the position update (:216-236)
  dfsph_cluster       the 33 lattice points within h of a centre, under a
                      converging velocity field: no density iteration, one
                      divergence iteration bounded by maxIterations (:512);
                      alpha of every point, the 20-neighbor cutoff
                      (:480-483), kappa_v = divErr alpha / (dt rho) (:531),
                      kv += divErr alpha (:534) and the velocity sweep
  dfsph_dt_*          one particle: the CFL rule with the square root of the
                      speed and its 0.1 floor, the 1.001 and 0.99 branches and
                      the clamps (:827-870)
  wcsph_cluster       4 fluid particles (densities above rho0) and one Akinci
                      boundary particle: the Tait pressure force
                      (src/WCSPHSolver.cpp:80-95,152-186) and XSPH with its
                      division by dt (src/XSPHViscosity.h:45-67)
  pcisph_cluster      the same particles and a fifth one below rest density,
                      one PCISPH iteration: predicted densities at x + v dt,
                      p += delta max(0, rho - rho0), the Spiky pressure
                      acceleration (src/PCISPHSolver.cpp:122-358, 389-421)
  surface_tension     3 particles, both cohesion branches: normals and
                      gamma K_ij (-m C r^ - (n_i - n_j))
                      (src/SurfaceTensionAkinci.h:46-96, src/Cohesion.h:29-40)
  adhesion            1 fluid and 2 boundary particles: -beta rho0 V_b A r^
                      (src/AdhesionForce.h:38-61, src/Adhesion.h:24-35) with
                      Akinci volumes 1 / sum W (src/AkinciBoundayModel.cpp:11-40)

Run: python tests/golden/make_solver_closed_forms.py
"""
import json
import os
from decimal import Decimal as D, getcontext

getcontext().prec = 40

R = D(1) / 64          # particle radius
H = 4 * R              # support radius
RHO0 = D(1000)
M = RHO0 * (2 * R) ** 3
EPS = D("1e-5")


def pi():
    getcontext().prec += 5

    def arctan_inv(x):
        x = D(x)
        term = 1 / x
        s = term
        x2 = x * x
        n, sign = 1, -1
        while True:
            term /= x2
            t = term / (2 * n + 1)
            if t == 0 or abs(t) < D(10) ** -(getcontext().prec + 2):
                break
            s += sign * t
            sign, n = -sign, n + 1
        return s

    p = 4 * (4 * arctan_inv(5) - arctan_inv(239))
    getcontext().prec -= 5
    return +p


PI = pi()


# ---------------------------------------------------------------- vectors
def add(a, b):
    return tuple(x + y for x, y in zip(a, b))


def sub(a, b):
    return tuple(x - y for x, y in zip(a, b))


def mul(s, a):
    return tuple(s * x for x in a)


def dot(a, b):
    return sum((x * y for x, y in zip(a, b)), D(0))


def norm(a):
    return dot(a, a).sqrt()


ZERO = (D(0), D(0), D(0))


def vsum(vs):
    out = ZERO
    for v in vs:
        out = add(out, v)
    return out


# ---------------------------------------------------------------- kernels
def W(rl):
    """Cubic spline (src/CubicSpline.h:29-48)."""
    q = rl / H
    k = 8 / (PI * H ** 3)
    if q <= D("0.5"):
        return k * (6 * q ** 3 - 6 * q ** 2 + 1)
    if q <= 1:
        return 2 * k * (1 - q) ** 3
    return D(0)


def gradW(r):
    """Cubic spline gradient (src/CubicSpline.h:55-73)."""
    rl = norm(r)
    q = rl / H
    if not (rl > EPS and q <= 1):
        return ZERO
    l = 48 / (PI * H ** 3)
    f = q * (3 * q - 2) if q <= D("0.5") else -(1 - q) ** 2
    return mul(l * f / (rl * H), r)


def spiky_grad(r):
    """src/Spiky.h:44-55."""
    rl = norm(r)
    if rl == 0 or rl > H:
        return ZERO
    return mul(-45 / (PI * H ** 6) * (H - rl) ** 2 / rl, r)


def poly6_grad(r):
    """src/Poly6.h:44-57."""
    r2 = dot(r, r)
    if r2 > H * H:
        return ZERO
    return mul(-(H * H - r2) ** 2 * 945 / (32 * PI * H ** 9), r)


def cohesion(rl):
    """src/Cohesion.h:29-40, the h^6/64 offset unscaled as written there."""
    c = 32 / (PI * H ** 9)
    if 2 * rl > H and rl <= H:
        return c * (H - rl) ** 3 * rl ** 3
    if rl > 0 and 2 * rl <= H:
        return c * 2 * (H - rl) ** 3 * rl ** 3 - H ** 6 / 64
    return D(0)


def adhesion(rl):
    """src/Adhesion.h:24-35 without its 0.007/h^3.25 coefficient."""
    if 2 * rl > H and rl <= H:
        return (-4 * rl * rl / H + 6 * rl - 2 * H).sqrt().sqrt()
    return D(0)


# ------------------------------------------------------------------ scenes
def dy(x):
    """A dyadic value given as a string fraction 'a/b' or a number."""
    if isinstance(x, str) and "/" in x:
        a, b = x.split("/")
        return D(a) / D(b)
    return D(str(x))


def vec(t):
    return tuple(dy(c) for c in t)


def out3(vs):
    return [[str(c) for c in v] for v in vs]


def out1(xs):
    return [str(x) for x in xs]


def akinci_volumes(xb):
    return [1 / sum((W(norm(sub(b, c))) for c in xb), D(0)) for b in xb]


def densities(x, xb, vb, at=None):
    """rho_i = sum_j m W_ij (self included) + sum_b rho0 V_b W_ib
    (src/SPHSolver.cpp:60-89); `at` replaces the fluid positions (PCISPH's
    predicted positions)."""
    p = x if at is None else at
    return [sum((M * W(norm(sub(p[i], p[j]))) for j in range(len(p))), D(0)) +
            sum((RHO0 * vb[b] * W(norm(sub(p[i], xb[b]))) for b in range(len(xb))), D(0))
            for i in range(len(p))]


def alphas(x, xb, vb, rho):
    out = []
    for i in range(len(x)):
        g = [mul(M, gradW(sub(x[i], x[j]))) for j in range(len(x))]
        s = vsum(g)
        s = add(s, vsum(mul(RHO0 * vb[b], gradW(sub(x[i], xb[b]))) for b in range(len(xb))))
        sq = sum((dot(v, v) for v in g), D(0))
        den = dot(s, s) + sq
        out.append(rho[i] / den if abs(den) > EPS else D(0))
    return out


def dfsph_pair(d, U, dt):
    """Two particles on the x axis approaching at +-U; one density iteration,
    no divergence iteration (two particles are below the 20-neighbor cutoff)."""
    c = D("0.5")
    x = [(c, c, c), (c + d, c, c)]
    v = [(U, D(1) / 8, D(0)), (-U, D(0), D(1) / 4)]
    far = [(D(1), D(1), D(1))]
    vb = akinci_volumes(far)
    rho = densities(x, far, vb)
    alpha = alphas(x, far, vb, rho)
    n = len(x)
    g = [[gradW(sub(x[i], x[j])) for j in range(n)] for i in range(n)]

    def pred(vv):
        out = []
        for i in range(n):
            dc = sum((M * dot(sub(vv[i], vv[j]), g[i][j]) for j in range(n)), D(0))
            out.append(max(rho[i] + dc * dt, RHO0))
        return out

    rs0 = pred(v)
    err = [rs - RHO0 for rs in rs0]
    kap = [err[i] * alpha[i] / (dt * dt * rho[i]) for i in range(n)]
    kacc = [err[i] * alpha[i] for i in range(n)]
    v1 = []
    for i in range(n):
        s = ZERO
        for j in range(n):
            ks = kap[i] + kap[j]
            if abs(ks) > EPS:
                s = add(s, mul(M * ks, g[i][j]))
        v1.append(sub(v[i], mul(dt, s)))
    rs1 = pred(v1)
    x1 = [add(x[i], mul(dt, v1[i])) for i in range(n)]
    assert min(rs0) > RHO0 + 100, rs0        # the density solve has work to do
    return {
        "r": str(R), "dt": str(dt), "x": out3(x), "v": out3(v), "boundary": out3(far),
        "expect": {"rho": out1(rho), "alpha": out1(alpha), "rhostar0": out1(rs0), "kappa": out1(kap),
                   "k": out1(kacc), "v": out3(v1), "rhostar": out1(rs1), "x": out3(x1)},
    }


def cluster_points():
    pts = []
    for i in range(-2, 3):
        for j in range(-2, 3):
            for k in range(-2, 3):
                if i * i + j * j + k * k <= 4:
                    pts.append((i, j, k))
    return pts


def dfsph_cluster(dt):
    """The 33 lattice points (spacing 2r) within h of a centre, moving with
    v = -64 (x - centre): the position update scales the cluster by 63/64
    (exact), then densities, alpha and the divergence error are formed at the
    new positions and one divergence iteration runs (the scene asks for at
    least 2, but the loop is bounded by maxIterations = 1, not by
    maxIterationsV = 5)."""
    c = D("0.5")
    s = 2 * R
    offs = cluster_points()
    x0 = [(c + i * s, c + j * s, c + k * s) for (i, j, k) in offs]
    rate = D(64)
    v = [mul(-rate, sub(p, (c, c, c))) for p in x0]
    assert rate * dt == D(1) / 64
    x = [add(x0[i], mul(dt, v[i])) for i in range(len(x0))]
    far = [(D(1), D(1), D(1))]
    vb = akinci_volumes(far)
    rho = densities(x, far, vb)
    alpha = alphas(x, far, vb, rho)
    n = len(x)
    g = [[gradW(sub(x[i], x[j])) for j in range(n)] for i in range(n)]

    def div(vv):
        out, cnt = [], []
        for i in range(n):
            de = sum((M * dot(sub(vv[i], vv[j]), g[i][j]) for j in range(n)), D(0))
            nn = sum(1 for j in range(n) if norm(g[i][j]) > 0)
            cnt.append(nn)
            out.append(D(0) if nn < 20 else max(de, D(0)))
        return out, cnt

    de0, cnt = div(v)
    kv = [de0[i] * alpha[i] / (dt * rho[i]) for i in range(n)]
    kvacc = [de0[i] * alpha[i] for i in range(n)]
    v1 = []
    for i in range(n):
        sacc = ZERO
        for j in range(n):
            ks = kv[i] + kv[j]
            if abs(ks) > EPS:
                sacc = add(sacc, mul(M * ks, g[i][j]))
        v1.append(sub(v[i], mul(dt, sacc)))
    de1, _ = div(v1)
    assert cnt[offs.index((0, 0, 0))] == 32 and min(cnt) < 20 and max(de0) > 0
    return {
        "r": str(R), "dt": str(dt), "x": out3(x0), "v": out3(v), "boundary": out3(far),
        "expect": {"x": out3(x), "rho": out1(rho), "alpha": out1(alpha), "diverr0": out1(de0),
                   "ngrad": cnt, "kappa_v": out1(kv), "kv": out1(kvacc), "v": out3(v1), "diverr": out1(de1)},
    }


def dfsph_dt(v, g, dt0, steps, min_it, cfl=D("0.4"), dt_max=D(1), dt_min=D("1e-6")):
    """Time steps of one free particle (no neighbors, alpha 0, no pressure):
    maxVel = max(0.1, sqrt(|v + (a + g) dt|)), cfl = c 0.4 2r / maxVel,
    dt' = cfl if cfl < dt, else 0.99 dt after more than 3 density iterations,
    else 1.001 dt; clamped.  The density loop runs exactly min_it iterations
    (its error stays 0: rho* is clamped to rho0)."""
    vel = v
    dt = dt0
    last_it = 0
    dts = []
    for _ in range(steps):
        cur = add(vel, mul(dt, g))
        mv = max(D("0.1"), norm(cur).sqrt())
        cf = cfl * D("0.4") * 2 * R / mv
        if cf < dt:
            nd = cf
        elif last_it > 3:
            nd = dt * D("0.99")
        else:
            nd = dt * D("1.001")
        nd = min(max(nd, dt_min), dt_max)
        dt = nd
        vel = add(vel, mul(dt, g))
        last_it = min_it
        dts.append(dt)
    return {"r": str(R), "v": [str(c) for c in v], "g": [str(c) for c in g], "dt0": str(dt0), "steps": steps,
            "min_iterations": min_it, "cfl": str(cfl), "dt_max": str(dt_max), "dt_min": str(dt_min),
            "expect": {"dt": out1(dts)}}


def tait(rho, kappa=D(1000), gam=7):
    return max(D(0), kappa * RHO0 / gam * ((rho / RHO0) ** gam - 1))


CL4 = [("0", "0", "0"), ("3/256", "0", "0"), ("0", "13/1024", "0"), ("0", "0", "7/512")]
CL4_V = [("1/8", "-1/16", "1/32"), ("-1/4", "1/8", "0"), ("0", "-1/8", "1/16"), ("1/16", "1/16", "-1/8")]
CL4_B = [("-1/64", "-1/128", "-3/256")]


def shifted(pts, c=D("0.5")):
    return [tuple(c + dy(a) for a in p) for p in pts]


def wcsph_cluster(dt, nu, nub):
    x = shifted(CL4)
    v = [vec(t) for t in CL4_V]
    xb = shifted(CL4_B)
    vb = akinci_volumes(xb)
    rho = densities(x, xb, vb)
    p = [tait(r) for r in rho]
    n = len(x)
    acc = []
    for i in range(n):
        a = ZERO
        # XSPH (src/XSPHViscosity.h:45-67): the fluid term over every
        # candidate, the boundary term with nu_b != 0
        for j in range(n):
            a = add(a, mul(nu * (M / rho[j]) * W(norm(sub(x[i], x[j]))) / dt, sub(v[j], v[i])))
        for b in range(len(xb)):
            a = add(a, mul(nub * (RHO0 * vb[b] / rho[i]) * W(norm(sub(x[i], xb[b]))) / dt, mul(-1, v[i])))
        # pressure (src/WCSPHSolver.cpp:152-186): p_i / rho_i^2 alone at the boundary
        for j in range(n):
            a = sub(a, mul(M * (p[i] / rho[i] ** 2 + p[j] / rho[j] ** 2), gradW(sub(x[i], x[j]))))
        for b in range(len(xb)):
            a = sub(a, mul(RHO0 * vb[b] * p[i] / rho[i] ** 2, gradW(sub(x[i], xb[b]))))
        acc.append(a)
    assert min(rho) > RHO0 + 10, rho
    return {"r": str(R), "dt": str(dt), "nu": str(nu), "nub": str(nub), "x": out3(x), "v": out3(v),
            "boundary": out3(xb), "expect": {"bvol": out1(vb), "rho": out1(rho), "p": out1(p), "a": out3(acc)}}


def pcisph_delta(dt):
    s = 2 * R
    sg, sq = ZERO, D(0)
    for i in (-1, 0, 1):
        for j in (-1, 0, 1):
            for k in (-1, 0, 1):
                gr = poly6_grad((-i * s, -j * s, -k * s))
                sg = add(sg, gr)
                sq += dot(gr, gr)
    beta = 2 * (M * dt / RHO0) ** 2
    return -1 / (beta * (-dot(sg, sg) - sq))


def pcisph_cluster(dt):
    x = shifted(CL4 + [("-3/64", "1/64", "0")])
    v = [vec(t) for t in CL4_V] + [vec(("1/2", "0", "-1/4"))]
    xb = shifted(CL4_B)
    vb = akinci_volumes(xb)
    n = len(x)
    delta = pcisph_delta(dt)
    # one iteration: v* = v (no forces, pacc = 0, g = 0), x* = x + v dt
    xs = [add(x[i], mul(dt, v[i])) for i in range(n)]
    rho = densities(x, xb, vb, at=xs)
    p = [delta * max(D(0), r - RHO0) for r in rho]
    pacc = []
    for i in range(n):
        a = ZERO
        for j in range(n):
            a = sub(a, mul(M * (p[i] / rho[i] ** 2 + p[j] / rho[j] ** 2), spiky_grad(sub(x[i], x[j]))))
        for b in range(len(xb)):
            a = sub(a, mul(RHO0 * vb[b] * p[i] / rho[i] ** 2, spiky_grad(sub(x[i], xb[b]))))
        pacc.append(a)
    v1 = [add(v[i], mul(dt, pacc[i])) for i in range(n)]
    x1 = [add(x[i], mul(dt, v1[i])) for i in range(n)]
    assert min(rho) < RHO0 < max(rho)
    return {"r": str(R), "dt": str(dt), "x": out3(x), "v": out3(v), "boundary": out3(xb),
            "expect": {"delta": str(delta), "rho": out1(rho), "p": out1(p), "pacc": out3(pacc), "v": out3(v1),
                       "x": out3(x1)}}


def surface_tension(gamma, dt):
    # particle 1 at q = 0.375 from particle 0 (inner branch), particle 2 at
    # q = 0.75 (outer branch); 1 and 2 are 0.0527 apart (q = 0.84)
    x = shifted([("0", "0", "0"), ("3/128", "0", "0"), ("0", "3/64", "0")])
    far = [(D(1), D(1), D(1))]
    vb = akinci_volumes(far)
    rho = densities(x, far, vb)
    n = len(x)
    nrm = [mul(H, vsum(mul(M / rho[j], gradW(sub(x[i], x[j]))) for j in range(n))) for i in range(n)]
    acc = []
    for i in range(n):
        a = ZERO
        for j in range(n):
            rij = sub(x[i], x[j])
            rl = norm(rij)
            if rl > 0 and rl <= H:
                kij = 2 * RHO0 / (rho[i] + rho[j])
                coh = mul(-M * cohesion(rl) / rl, rij)
                cur = mul(-1, sub(nrm[i], nrm[j]))
                a = add(a, mul(gamma * kij, add(coh, cur)))
        acc.append(a)
    assert max(rho) < RHO0     # no pressure
    return {"r": str(R), "dt": str(dt), "gamma": str(gamma), "x": out3(x), "boundary": out3(far),
            "expect": {"rho": out1(rho), "normal": out3(nrm), "a": out3(acc)}}


def adhesion_case(beta, dt):
    x = shifted([("0", "0", "0")])
    xb = shifted([("5/128", "0", "0"), ("0", "0", "-3/64")])
    vb = akinci_volumes(xb)
    rho = densities(x, xb, vb)
    a = ZERO
    for b in range(len(xb)):
        rib = sub(x[0], xb[b])
        rl = norm(rib)
        a = sub(a, mul(beta * RHO0 * vb[b] * adhesion(rl) / rl, rib))
    assert rho[0] < RHO0
    return {"r": str(R), "dt": str(dt), "beta": str(beta), "x": out3(x), "boundary": out3(xb),
            "expect": {"bvol": out1(vb), "rho": out1(rho), "a": out3([a])}}


def main():
    dt = D(1) / 4096
    out = {
        "dfsph_pair_q0.1875": dfsph_pair(D(3) / 256, D(256), dt),
        "dfsph_pair_q0.6875": dfsph_pair(D(11) / 256, D(768), dt),
        "dfsph_cluster": dfsph_cluster(dt),
        # |v| = 100: sqrt(|v|) = 10 puts the CFL step below dt
        "dfsph_dt_cfl": dfsph_dt(vec(("60", "80", "0")), ZERO, D("1e-3"), 1, 2),
        # at rest under g = 2500: sqrt(|g dt|) = sqrt(10)
        "dfsph_dt_gravity": dfsph_dt(ZERO, vec(("0", "0", "-2500")), D("4e-3"), 1, 2),
        # |v| = 3.25: its square root puts the CFL step (2.77e-3) above dt: 1.001 dt
        "dfsph_dt_sqrt": dfsph_dt(vec(("3/4", "1", "3")), ZERO, D("2e-3"), 1, 2),
        # at rest: the 0.1 floor; 5 iterations in step 1 make step 2 take 0.99
        "dfsph_dt_099": dfsph_dt(ZERO, ZERO, D("1e-3"), 3, 5),
        # 3 iterations are not more than 3: 1.001 again; the maximum clamps
        "dfsph_dt_1001_clamp": dfsph_dt(ZERO, ZERO, D("1e-3"), 3, 3, dt_max=D("1.0015e-3")),
        "wcsph_cluster": wcsph_cluster(D("5e-4"), D("0.01"), D("0.02")),
        "pcisph_cluster": pcisph_cluster(D(1) / 4096),
        "surface_tension": surface_tension(D("0.05"), D("5e-4")),
        "adhesion": adhesion_case(D("0.1"), D("5e-4")),
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "solver_closed_forms.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", path)


if __name__ == "__main__":
    main()
