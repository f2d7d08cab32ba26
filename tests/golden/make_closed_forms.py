"""Closed-form golden values for the SPH kernels (no reference run needed).

Writes tests/golden/closed_forms.json.  Every value follows from the kernel
formulas of the reference (src/CubicSpline.h:29-48, src/Poly6.h:44-57,
src/PCISPHSolver.cpp:389-421) evaluated on exact lattices with 40-digit
decimal arithmetic, independently of the oracle's code:

  lattice_density_ratio   rho / rho0 of an interior particle of a simple
                          cubic lattice at spacing h/2 (m = rho0 (h/2)^3):
                          sum_n W(|n| h/2) (h/2)^3 = sum / pi
  wall_volume_over_h3     Akinci volume of a particle inside an infinite flat
                          wall sampled at spacing h/2: V = 1 / sum W
  pcisph_delta_h2_dt2     PCISPH delta * dt^2 / h^2 with Poly6 gradients
                          on the 3x3x3 prototype at spacing h/2
  lattice_counts          candidates with q <= 1, nonzero W, nonzero gradW
"""
import json
import os
from decimal import Decimal, getcontext

getcontext().prec = 40


def pi():
    # Machin-like series for pi (40 digits)
    getcontext().prec += 5

    def arctan_inv(x):
        x = Decimal(x)
        term = 1 / x
        s = term
        x2 = x * x
        n = 1
        sign = -1
        while True:
            term /= x2
            t = term / (2 * n + 1)
            if t == 0 or abs(t) < Decimal(10) ** -(getcontext().prec + 2):
                break
            s += sign * t
            sign = -sign
            n += 1
        return s

    p = 4 * (4 * arctan_inv(5) - arctan_inv(239))
    getcontext().prec -= 5
    return +p


def wk(q):
    """Cubic spline W / k as a function of q = r/h."""
    if q <= Decimal("0.5"):
        return 6 * q ** 3 - 6 * q ** 2 + 1
    if q <= 1:
        return 2 * (1 - q) ** 3
    return Decimal(0)


def main():
    PI = pi()
    half = Decimal("0.5")
    # simple cubic lattice at spacing h/2: q = |n| / 2
    s3 = Decimal(0)
    n_cand = n_w = n_g = 0
    for i in range(-2, 3):
        for j in range(-2, 3):
            for k in range(-2, 3):
                q = (Decimal(i * i + j * j + k * k)).sqrt() * half
                if q <= 1:
                    n_cand += 1
                    w = wk(q)
                    s3 += w
                    n_w += 1 if w != 0 else 0
                    g = (q * (3 * q - 2)) if q <= half else (1 - q) ** 2
                    n_g += 1 if (q > 0 and g != 0) else 0
    # 2D square lattice (flat wall) at spacing h/2
    s2 = Decimal(0)
    for i in range(-2, 3):
        for j in range(-2, 3):
            q = (Decimal(i * i + j * j)).sqrt() * half
            if q <= 1:
                s2 += wk(q)
    k_over = 8 / PI          # k h^3
    lattice_ratio = s3 / PI  # sum W (h/2)^3 = (8/pi) sum (1/8)
    wall_v = 1 / (k_over * s2)
    # PCISPH: sum |grad Poly6|^2 over the prototype, in units of c^2 h^10
    sq = Decimal(0)
    for i in (-1, 0, 1):
        for j in (-1, 0, 1):
            for k in (-1, 0, 1):
                r2 = Decimal(i * i + j * j + k * k) / 4
                if r2 <= 1:
                    sq += (1 - r2) ** 4 * r2
    # delta = 1/(beta sum|grad|^2), beta = h^6 dt^2 / 32, c = 945/(32 pi h^9)
    c2 = (Decimal(945) / (32 * PI)) ** 2
    delta_h2_dt2 = 32 / (c2 * sq)
    out = {
        "lattice_sum_W_over_k": str(s3),
        "lattice_density_ratio": str(lattice_ratio),
        "wall_sum_W_over_k": str(s2),
        "wall_volume_over_h3": str(wall_v),
        "pcisph_sum_sq_grad_units": str(sq),
        "pcisph_delta_h2_dt2": str(delta_h2_dt2),
        "lattice_counts": {"candidates_q_le_1": n_cand, "nonzero_W": n_w, "nonzero_gradW": n_g},
        "cubic_W0_h3": str(8 / PI),
        "box_shell_counts": {
            "3x1.5x1.5@h0.1": 9002, "3x1.7x1.7@h0.1": 10474, "4x2x2@h0.04": 100002,
            "4x2.02x2.02@h0.04": 101204, "0.5x0.5x0.1@h0.02": 7002, "2x2x2@h0.02": 240002,
        },
    }
    # shell counts: (nx+1)(ny+1)(nz+1) - (nx-1)(ny-1)(nz-1) with n = ceil(L / (h/2))
    def shell(L, h):
        import math
        n = [int(math.ceil(Decimal(str(a)) / (Decimal(str(h)) / 2))) for a in L]
        return (n[0] + 1) * (n[1] + 1) * (n[2] + 1) - (n[0] - 1) * (n[1] - 1) * (n[2] - 1)
    for key, (L, h) in {"3x1.5x1.5@h0.1": ((3, 1.5, 1.5), 0.1), "3x1.7x1.7@h0.1": ((3, 1.7, 1.7), 0.1),
                        "4x2x2@h0.04": ((4, 2, 2), 0.04), "4x2.02x2.02@h0.04": ((4, 2.02, 2.02), 0.04),
                        "0.5x0.5x0.1@h0.02": ((0.5, 0.5, 0.1), 0.02), "2x2x2@h0.02": ((2, 2, 2), 0.02)}.items():
        assert shell(L, h) == out["box_shell_counts"][key], key
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "closed_forms.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
