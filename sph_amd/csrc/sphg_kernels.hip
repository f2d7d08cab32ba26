// sphg_kernels.hip -- gfx950 kernels of the SPH solver step.
//
// Every kernel takes a scene table and a chunk table; block b works on
// chunk b (kBlock = 256 particles or kCellChunk = 1024 cells of one
// scene).  Sums of a scene are formed in a fixed order: a butterfly inside
// each wavefront, wave results combined in index order, one partial per
// chunk, and then the partials added in chunk order by k_finalize (one
// block per scene, launched after the sweep).  Results therefore do not
// depend on scheduling.
//
// Reference algorithm: josemanuel28v/SPH src/SPHSolver.cpp,
// src/DFSPHSolver.cpp, src/WCSPHSolver.cpp, src/PCISPHSolver.cpp, the kernel
// headers and the force headers; each kernel names the lines it restates.

#include "sphg_internal.h"
#include "sphg_kernels.h"

#include <type_traits>

namespace sphg {

// ------------------------------------------------------------ utilities

// f(integral_constant<int, I>) for I = 0 .. N-1, unrolled at compile time
template <int N, int I = 0, class F>
__device__ inline void static_for(F&& f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<N, I + 1>(f);
    }
}

// Global-memory access.  The arrays come out of the scene table, so the
// compiler cannot prove their address space; these casts make it emit
// global (not flat) loads and stores.
typedef float v4f __attribute__((ext_vector_type(4)));
template <class T> using gptr = __attribute__((address_space(1))) T*;
__device__ inline float4 ld4(const float4* p, size_t i) {
    const v4f t = ((gptr<const v4f>)p)[i];
    return make_float4(t.x, t.y, t.z, t.w);
}
__device__ inline void st4(float4* p, size_t i, float4 v) {
    v4f t;
    t.x = v.x; t.y = v.y; t.z = v.z; t.w = v.w;
    ((gptr<v4f>)p)[i] = t;
}
template <class T> __device__ inline T ldg(const T* p, size_t i) { return ((gptr<const T>)p)[i]; }
// p as a global-address-space pointer, for p[i] and p->field of scalar types
template <class T> __device__ inline gptr<T> G(T* p) { return (gptr<T>)p; }
// Element j of a float4 array of fewer than 2^28 elements: the 32-bit byte
// offset lets the load take the uniform-base + 32-bit-offset address form
// (no 64-bit address arithmetic per element).
__device__ inline float4 ld4o(const float4* p, uint32_t j) {
    const v4f t = *(gptr<const v4f>)((gptr<const char>)p + (j << 4));
    return make_float4(t.x, t.y, t.z, t.w);
}
typedef float v2f __attribute__((ext_vector_type(2)));
__device__ inline float2 ld2o(const float2* p, uint32_t j) {
    const v2f t = *(gptr<const v2f>)((gptr<const char>)p + (j << 3));
    return make_float2(t.x, t.y);
}
// The xyz of element j of a float4 array: a 12-byte load.  (Measured on
// gfx950: slightly faster for the list build's survivors, slower than the
// 16-byte load for the solver sweeps' gathers.)
typedef float v3f __attribute__((ext_vector_type(3)));
__device__ inline float4 ld3o(const float4* p, uint32_t j) {
    const v3f t = *(gptr<const v3f>)((gptr<const char>)p + (j << 4));
    return make_float4(t.x, t.y, t.z, 0.0f);
}
template <class T> __device__ inline void stg(T* p, size_t i, T v) { ((gptr<T>)p)[i] = v; }
template <class T> __device__ inline void stg(T* p, uint32_t i, T v) {
    *(gptr<T>)((gptr<char>)p + (size_t)(i * (uint32_t)sizeof(T))) = v;
}


__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}
// Result valid in every thread; the wave results are added in wave order.
// sh: kWaves doubles of shared memory.
__device__ inline double block_sum(double v, double* sh) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double r = sh[0];
#pragma unroll
    for (int k = 1; k < kWaves; ++k) r += sh[k];
    __syncthreads();
    return r;
}
__device__ inline double block_max(double v, double* sh) {
    v = wave_max(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double r = sh[0];
#pragma unroll
    for (int k = 1; k < kWaves; ++k) r = fmax(r, sh[k]);
    __syncthreads();
    return r;
}

// Sum (or max) of one scene's chunk partials in chunk order.
// Each thread combines partials t, t + kBlock, t + 2 kBlock, ... in that order; the
// loads of 4096 / kBlock of them are issued before the first is used, so
// the loop waits on memory once for up to 4096 partials (a 1M-particle
// scene's chunks) instead of once per partial.
constexpr int kPartU = 4096 / kBlock;
__device__ inline double scene_partial_sum(const double* partial, int n, double* sh) {
    double acc = 0.0;
    for (int b = threadIdx.x; b < n; b += kBlock * kPartU) {
        double v[kPartU];
#pragma unroll
        for (int u = 0; u < kPartU; ++u) v[u] = (b + u * kBlock < n) ? ldg(partial, (size_t)(b + u * kBlock)) : 0.0;
#pragma unroll
        for (int u = 0; u < kPartU; ++u) acc += v[u];
    }
    return block_sum(acc, sh);
}
__device__ inline double scene_partial_max(const double* partial, int n, double* sh) {
    double acc = -1e300;
    for (int b = threadIdx.x; b < n; b += kBlock * kPartU) {
        double v[kPartU];
#pragma unroll
        for (int u = 0; u < kPartU; ++u) v[u] = (b + u * kBlock < n) ? ldg(partial, (size_t)(b + u * kBlock)) : -1e300;
#pragma unroll
        for (int u = 0; u < kPartU; ++u) acc = fmax(acc, v[u]);
    }
    return block_max(acc, sh);
}

// ------------------------------------------------ SPH kernels (FP64 math)
//
// State is FP32 in HBM; every neighbor sum is formed in FP64.  On CDNA3/4
// an FP64 FMA issues at the rate of an unpacked FP32 FMA, and the FP32 sums
// lose the small differences the solvers run on (rho* - rho0 of a density
// near 1000).  1/|r| is an FP32 rsqrt refined by one FP64 Newton step
// (relative error ~1e-14), so no FP64 transcendental is issued.

// 1/sqrt(r2), zero for r2 <= 1e-30 (coincident points contribute no gradient).
// The seed is the hardware FP32 reciprocal square root: every r2 it is kept
// for is a normal float (1e-30 > 2^-126), where rsqrtf is that same
// instruction without its subnormal rescaling.
__device__ inline double rsqrt_d(double r2) {
    const double y = (double)__builtin_amdgcn_rsqf((float)r2);
    const double z = y * (1.5 - 0.5 * r2 * y * y);
    return (r2 > 1e-30) ? z : 0.0;
}
__device__ inline double sel(bool c, double a, double b) { return c ? a : b; }

struct Pair {
    double dx, dy, dz, r2, rl, inv, q;
};
__device__ inline Pair pair_of(double xi, double yi, double zi, float4 b, const KCoefD& k) {
    Pair p;
    p.dx = xi - (double)b.x;
    p.dy = yi - (double)b.y;
    p.dz = zi - (double)b.z;
    p.r2 = p.dx * p.dx + p.dy * p.dy + p.dz * p.dz;
    p.inv = rsqrt_d(p.r2);
    p.rl = p.r2 * p.inv;
    p.q = p.rl * k.invh;
    return p;
}
// Cubic spline W (src/CubicSpline.h:29-48); W_in for q <= 1.
__device__ inline double W_in(double q, const KCoefD& k) {
    const double om = 1.0 - q;
    const double w1 = k.ck * ((6.0 * q - 6.0) * q * q + 1.0);
    const double w2 = 2.0 * k.ck * om * om * om;
    return sel(q <= 0.5, w1, w2);
}
__device__ inline double W_d(double q, const KCoefD& k) { return sel(q <= 1.0, W_in(q, k), 0.0); }
// Cubic spline gradient factor: gradW(r) = g r, zero for |r| <= 1e-5 or
// q > 1 (src/CubicSpline.h:55-73).  For q <= 1/2 the reference's
// l q (3q-2) / (|r| h) equals l (3q-2) / h^2.  G_in for |r| > 1e-5, q <= 1.
__device__ inline double G_in(const Pair& p, const KCoefD& k) {
    const double om = 1.0 - p.q;
    const double g1 = k.clh2 * (3.0 * p.q - 2.0);
    const double g2 = -k.clh * om * om * p.inv;
    return sel(p.q <= 0.5, g1, g2);
}
__device__ inline double G_d(const Pair& p, const KCoefD& k) {
    return sel(p.rl > 1e-5 && p.q <= 1.0, G_in(p, k), 0.0);
}
// Cubic gradient factor of a neighbor-list member for the solver sweeps.
// List members have q <= 1 by construction and the padding entries are the
// particle itself (r = 0), so the q <= 1 test of G_d is dropped, and 1/r is
// the raw FP32 rsqrt refined by one FP64 Newton step with no guard: r = 0
// gives rl = NaN, which fails rl > 1e-5 and selects 0, as G_d does.
struct PairG {
    double dx, dy, dz, g;
};
__device__ inline PairG pairg_of(double xi, double yi, double zi, float4 b, const KCoefD& k) {
    PairG p;
    p.dx = xi - (double)b.x;
    p.dy = yi - (double)b.y;
    p.dz = zi - (double)b.z;
    const double r2 = p.dx * p.dx + p.dy * p.dy + p.dz * p.dz;
    const double y = (double)__builtin_amdgcn_rsqf((float)r2);
    const double e = fma(-r2 * y, y, 1.0);          // 1 - r2 y^2
    const double inv = fma(0.5 * y, e, y);          // one Newton step
    const double rl = r2 * inv;
    // with q = r/h: g = clh2 (3q - 2) = gA r + gB for q <= 1/2, else
    // -clh (1 - q)^2 / r
    const double om = fma(-k.invh, rl, 1.0);
    const double g1 = fma(k.gA, rl, k.gB);
    const double g2 = (-k.clh * inv) * (om * om);
    p.g = sel(rl > 1e-5, sel(rl <= k.hHalf, g1, g2), 0.0);
    return p;
}

// Spiky gradient factor (src/Spiky.h:44-55): -45/(pi h^6) (h-r)^2 / r.
__device__ inline double Gs_d(const Pair& p, const KCoefD& k) {
    const double d = k.h - p.rl;
    const double g = -k.spg * d * d * p.inv;
    return sel(p.r2 > 0.0 && p.rl <= k.h, g, 0.0);
}
// Cohesion C(r) (src/Cohesion.h:29-40) for 0 < r <= h; `outer` is the
// branch test 2r > h.  As written there the h^6/64 offset is not scaled by
// 32/(pi h^9), so C jumps at r = h/2 and the branch is decided exactly by
// the caller near that point.
__device__ inline double C_d(double rl, bool outer, const KCoefD& k) {
    const double d = k.h - rl;
    const double t = d * d * d * rl * rl * rl;
    return outer ? k.co1 * t : k.co1 * 2.0 * t - k.co2;
}
// Adhesion A(r) (src/Adhesion.h:24-35): (-4r^2/h + 6r - 2h)^(1/4) on
// (h/2, h], evaluated as ((2/h)(2r-h)(h-r))^(1/4), the same polynomial in a
// form that cannot round below zero.
__device__ inline double A_d(double rl, const KCoefD& k) {
    if (!(2.0 * rl > k.h && rl <= k.h)) return 0.0;
    const double v = (2.0 * k.invh) * (2.0 * rl - k.h) * (k.h - rl);
    return sqrt(sqrt(fmax(v, 0.0)));
}

// Exact FP64 distance test, operation for operation the reference's
// length(ri - rj) / h on FP32 positions promoted to FP64
// (src/CubicSpline.h:55-60, glm dot summed left to right).
struct Exact {
    double rl, q;
};
__device__ inline Exact exact_q(float4 a, float4 b, double h) {
    const double dx = __dsub_rn((double)a.x, (double)b.x);
    const double dy = __dsub_rn((double)a.y, (double)b.y);
    const double dz = __dsub_rn((double)a.z, (double)b.z);
    const double d2 = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
    Exact e;
    e.rl = __dsqrt_rn(d2);
    e.q = __ddiv_rn(e.rl, h);
    return e;
}
// Whether the reference's FP64 gradient is nonzero (its neighbor count of
// src/DFSPHSolver.cpp:446-447).
__device__ inline bool grad_nonzero64(const Exact& e) {
    if (!(e.rl > 1.0e-5 && e.q <= 1.0)) return false;
    if (e.q <= 0.5) return __dmul_rn(e.q, __dsub_rn(__dmul_rn(3.0, e.q), 2.0)) != 0.0;
    const double f = __dsub_rn(1.0, e.q);
    return __dmul_rn(f, f) != 0.0;
}

// List layout inside a 64-particle group: blocks of 4 slots, lane-major
// within a block, so slot s of lane l is element (s/4 * 64 + l) * 4 + s%4
// of the group's maxNbr * 64 entries, and a lane reads its next 4 slots
// with one 16-byte load (a wavefront reads 1 KB contiguous).
__device__ inline uint32_t slot_of(int s) { return (uint32_t)(((s >> 2) * kGroup) << 2) + (uint32_t)(s & 3); }
// Element offset of lane (i & 63)'s first slot (maxNbr is a multiple of 4).
__device__ inline uint32_t lane_base(const DevScene& S, int i) {
    return (uint32_t)(i >> 6) * (uint32_t)S.maxNbr * (uint32_t)kGroup + ((uint32_t)(i & 63) << 2);
}
__device__ inline const uint32_t* list_of(const DevScene& S, int i) { return S.nbr + lane_base(S, i); }
typedef uint32_t v4u __attribute__((ext_vector_type(4)));
// slots 4b .. 4b+3 of a lane
__device__ inline v4u ld_block(const uint32_t* L, int b) { return ((gptr<const v4u>)L)[(size_t)b * kGroup]; }
// The group's padded list length, uniform across the wavefront (a
// wavefront is one 64-particle group: chunks start at multiples of 256).
__device__ inline int group_count(const DevScene& S, int i) {
    return __builtin_amdgcn_readfirstlane((int)ldg(S.gcnt, (size_t)(i >> 6)));
}

// Visit particle i's padded neighbor list four entries at a time: the four
// indices are loaded, then the four neighbors' data (load(j)), then the
// four terms are evaluated (body(j, data)), so every lane has several
// loads in flight.  The trip count is uniform across the wavefront.
template <class Load, class Body>
__device__ inline void sweep4(const DevScene& S, int i, Load&& load, Body&& body) {
    const uint32_t* L = list_of(S, i);
    const int gc = group_count(S, i);
    if (gc == 0) return;
    // the indices of the next 4 entries are requested before the neighbor
    // data of the current 4 is used, so the list stream (HBM) and the
    // neighbor loads (mostly L2) overlap instead of forming one chain
    v4u jn = ld_block(L, 0);
    for (int q = 0; q < gc; q += 4) {
        const uint32_t jj[4] = {jn.x, jn.y, jn.z, jn.w};
        decltype(load(0u)) dd[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) dd[u] = load(jj[u]);
        // unconditional (the last pass re-reads its own entries), so the
        // wait for this pass's data leaves the 4 index loads in flight
        const int qn = (q + 4 < gc) ? q + 4 : q;
        jn = ld_block(L, qn >> 2);
        // keeps the compiler from sinking these loads into the next pass
        asm volatile("" ::: "memory");
#pragma unroll
        for (int u = 0; u < 4; ++u) body(jj[u], dd[u]);
    }
}

struct XV {
    float4 x, v;
};
// x (with v_x in w) and (v_y, v_z) of a neighbor: 24 B, against 32 for x and v
struct XVP {
    float4 xv;
    float2 yz;
};
// the density solve's first error sweep: x (with u_x in w), v0 and (u_y, u_z)
struct XVU {
    float4 x, v;
    float2 yz;
};
struct NbF {
    float4 x, v, n;
    float d;
};
__device__ inline uint32_t* list_of_w(const DevScene& S, int i) { return S.nbr + lane_base(S, i); }

// Linear cell of a position: key floor(x/h) in FP64 (src/HashTable.h:81),
// clamped into the padded grid.  Clamping keeps any two particles closer
// than h in adjacent cells, so the 27-cell walk still sees every neighbor.
__device__ inline int cell_of(const DevScene& S, float4 p) {
    double fx = floor(__ddiv_rn((double)p.x, S.hD));
    double fy = floor(__ddiv_rn((double)p.y, S.hD));
    double fz = floor(__ddiv_rn((double)p.z, S.hD));
    fx = (fx == fx) ? fx : (double)S.kx0;
    fy = (fy == fy) ? fy : (double)S.ky0;
    fz = (fz == fz) ? fz : (double)S.kz0;
    const int ix = (int)fmin(fmax(fx - (double)S.kx0 + 1.0, 1.0), (double)(S.nx - 2));
    const int iy = (int)fmin(fmax(fy - (double)S.ky0 + 1.0, 1.0), (double)(S.ny - 2));
    const int iz = (int)fmin(fmax(fz - (double)S.kz0 + 1.0, 1.0), (double)(S.nz - 2));
    return ix + S.nx * (iy + S.ny * iz);
}

// Walk the 27 cells around cell c as 9 x-rows of 3 contiguous cells.
template <class F>
__device__ inline void walk27(const int* __restrict__ start, int c, int nx, int nxny, F&& f) {
#pragma unroll 1
    for (int dz = -1; dz <= 1; ++dz) {
#pragma unroll 1
        for (int dy = -1; dy <= 1; ++dy) {
            const int base = c + dy * nx + dz * nxny;
            const int s = start[base - 1];
            const int e = start[base + 2];
            for (int j = s; j < e; ++j) f(j);
        }
    }
}

// walk27 with the data of four candidates loaded before any is visited
// (`load(j)` then `visit(j, data)`, in walk order), so a lane has four
// independent loads in flight instead of one chain.
template <class Load, class Visit>
__device__ inline void walk27x4(const int* __restrict__ start, int c, int nx, int nxny, Load&& load, Visit&& visit) {
#pragma unroll 1
    for (int dz = -1; dz <= 1; ++dz) {
#pragma unroll 1
        for (int dy = -1; dy <= 1; ++dy) {
            const int base = c + dy * nx + dz * nxny;
            const int s = start[base - 1];
            const int e = start[base + 2];
            for (int j0 = s; j0 < e; j0 += 4) {
                decltype(load(0)) d[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) d[u] = load(min(j0 + u, e - 1));
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (j0 + u < e) visit(j0 + u, d[u]);
            }
        }
    }
}

// PCISPH shell width: 2.5 times the largest predicted displacement of the
// previous step (so a step's displacements may grow by 25%), at least
// 0.05 h.  The same for every kernel of a step (the previous step's maximum
// is copied at the step's start).
__device__ inline float pci_skin(const DevScene& S) {
    const float md = sqrtf(__uint_as_float(S.ss->maxD2Prev));
    return fmaxf(0.05f * S.kc.h, 2.5f * md);
}
// A predicted displacement longer than half the skin lets two particles
// outside each other's list and shell come within h, so the step's
// predicted densities take the 27-cell walk.
__device__ inline void shell_guard(const DevScene& S, double dx, double dy, double dz) {
    if (!S.pciShell) return;
    const double half = 0.5 * (double)pci_skin(S);
    if (dx * dx + dy * dy + dz * dz > half * half) atomicOr(&S.ss->shellBad, 1u);
}

// Poly6 W (src/Poly6.h:31-42) of a squared distance.
__device__ inline double P6_d(double r2, const KCoefD& k) {
    const double d = k.h2 - r2;
    return sel(r2 <= k.h2, k.p6v * d * d * d, 0.0);
}

// The collision projection of PCISPHBoundaryModel (src/PCISPHBoundaryModel.
// cpp:219-309, 385-439) for one fluid particle at position (px, py, pz):
// boundary candidates of the position's cell (none when no particle was
// ever inserted in that cell: the reference's hash table has no entry for
// it) with |x - r_b| < h/2 weigh w = max(0, (h/2 - |q_b|) / (h/2)), q_b =
// (qx, qy, qz) - r_b; returns n = glm::normalize(sum w n_b) (NaN for a zero
// sum) and the shift (1 / sum w)(sum w (h/2 - |q_b|)), applied when sum w > 0.
struct PbHit {
    double nx, ny, nz, s;
    bool hit;
};
__device__ inline PbHit pb_project(const DevScene& S, double px, double py, double pz, float4 pf, double qx, double qy,
                                   double qz) {
    const KCoefD k = S.kd;
    const double R = 0.5 * k.h;
    double ax = 0.0, ay = 0.0, az = 0.0, ws = 0.0, w2 = 0.0;
    const int cl = cell_of(S, pf);
    if (S.everOcc[cl]) {
        walk27(S.bCellStart, cl, S.nx, S.nx * S.ny, [&](int b) {
            const float4 y = S.xb[b];
            const double dx = px - (double)y.x, dy = py - (double)y.y, dz = pz - (double)y.z;
            if (sqrt(dx * dx + dy * dy + dz * dz) < R) {
                const double ex = qx - (double)y.x, ey = qy - (double)y.y, ez = qz - (double)y.z;
                const double lq = sqrt(ex * ex + ey * ey + ez * ez);
                const double w = fmax(0.0, (R - lq) / R);
                ws += w;
                w2 += w * (R - lq);
                const float4 nb = S.bn[b];
                ax += w * (double)nb.x;
                ay += w * (double)nb.y;
                az += w * (double)nb.z;
            }
        });
    }
    const double inv = 1.0 / sqrt(ax * ax + ay * ay + az * az);
    PbHit h;
    h.nx = ax * inv;
    h.ny = ay * inv;
    h.nz = az * inv;
    h.hit = ws > 0.0;
    h.s = h.hit ? (1.0 / ws) * w2 : 0.0;
    return h;
}
// correctVelocities (src/PCISPHBoundaryModel.cpp:311-383): where |n| > 0
// (false for NaN), v = tangentialFct (u - (u.n) n) - normalFct (u.n) n.
__device__ inline bool pb_velocity(const DevScene& S, double nx, double ny, double nz, double& ux, double& uy,
                                   double& uz) {
    if (!(sqrt(nx * nx + ny * ny + nz * nz) > 0.0)) return false;
    const double d = ux * nx + uy * ny + uz * nz;
    const double nvx = d * nx, nvy = d * ny, nvz = d * nz;
    ux = S.pbTf * (ux - nvx) - S.pbNf * nvx;
    uy = S.pbTf * (uy - nvy) - S.pbNf * nvy;
    uz = S.pbTf * (uz - nvz) - S.pbNf * nvz;
    return true;
}


// ================================================================ grid

// Cell key + histogram (src/HashTable.h:79-86 insertion, as a count).
// Adds one to counter[key] for every lane with key >= 0 and returns the
// counter's value before this lane's unit, a slot unique to the lane.  The
// particles of a wavefront are in the previous step's cell order and most
// keep their cell, so runs of consecutive lanes share a key: each run adds
// its length with one atomic (a dense fluid has ~8 lanes per run), and its
// lanes take consecutive slots.  Every lane of the wavefront calls it; lanes
// with nothing to count pass -1 and must all come after the counting ones.
__device__ inline int wave_count_add(int* counter, int key) {
    const int lane = threadIdx.x & 63;
    const int prev = __shfl_up(key, 1);
    const bool head = (lane == 0) || (key != prev);
    const uint64_t heads = __ballot(head);
    const uint64_t upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);   // lanes 0 .. lane
    const int first = 63 - __clzll((long long)(heads & upto));             // this lane's run head
    const uint64_t later = heads & ~upto;
    const int end = later ? __ffsll((long long)later) - 1 : 64;             // one past the run
    int base = 0;
    if (head && key >= 0)
        base = __hip_atomic_fetch_add(G(counter) + key, end - lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    base = __shfl(base, first);
    return base + (lane - first);
}

// The start of a WCSPH or PCISPH step: the scene's step counters, and for
// PCISPH zero pressures and pressure accelerations and the pressure loop's
// start (src/PCISPHSolver.cpp:21-50); `active` counts the scenes whose loop
// runs.
__device__ inline void step_begin(const DevScene& S, const Chunk& c, int i, int* active) {
    if (S.method == PCISPH && i < S.n) {
        S.p[i] = 0.0f;
        S.pacc[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    if (c.begin == 0 && threadIdx.x == 0) {
        Scalars* ss = S.ss;
        ss->steps += 1u;
        ss->shellBad = 0u;
        ss->maxD2Prev = ss->maxD2;
        ss->maxD2 = 0u;
        if (S.method == PCISPH) {
            ss->it = 0;
            const bool run = (ss->maxDensErr > S.eta || 0 < S.minIt) && 0 < S.maxIt;
            ss->done = run ? 0 : 1;
            if (run) atomicAdd(active, 1);
        }
    }
}

// Cell key of each particle and its count in its cell; with BEGIN also the
// start of a WCSPH or PCISPH step (one launch fewer per step).
template <bool BEGIN>
__global__ __launch_bounds__(kBlock) void k_keys(const DevScene* __restrict__ T, const Chunk* __restrict__ C,
                                                 int* active) {
    const Chunk c = C[blockIdx.x];
    const DevScene& S = T[c.scene];
    const int i = c.begin + threadIdx.x;
    if (BEGIN) step_begin(S, c, i, active);
    // the lists about to be built start without an overflow
    if (c.begin == 0 && threadIdx.x == 0) G(S.ss)->listOver = 0u;
    int cl = -1;
    if (i < S.n) {
        cl = cell_of(S, ld4(S.x, (size_t)i));
        G(S.cellU)[i] = cl;
    }
    (void)wave_count_add(S.cellCount, cl);
}

// Per cell chunk: total count.
__global__ __launch_bounds__(kBlock) void k_scan_sum(const DevScene* __restrict__ T, const Chunk* __restrict__ C) {
    const Chunk c = C[blockIdx.x];
    const DevScene& S = T[c.scene];
    __shared__ double sh[kWaves];
    int acc = 0;
    const int c0 = c.begin + threadIdx.x * kCellPerThread;
#pragma unroll
    for (int u = 0; u < kCellPerThread; ++u) {
        const int cc = c0 + u;
        if (cc < S.ncells) acc += ldg(S.cellCount, (size_t)cc);
    }
    const double tot = block_sum((double)acc, sh);
    if (threadIdx.x == 0) G(S.cellChunkSum)[c.begin / kCellChunk] = (int)tot;
}

__device__ inline int wave_incl_scan(int v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int t = __shfl_up(v, o);
        if (lane >= o) v += t;
    }
    return v;
}

// One block per scene: exclusive scan of the chunk totals, in place, kBlock
// totals per pass with the running total carried between passes.
__global__ __launch_bounds__(kBlock) void k_scan_offsets(const DevScene* __restrict__ T, int nscenes) {
    const int s = blockIdx.x;
    if (s >= nscenes) return;
    const DevScene& S = T[s];
    __shared__ int wsum[kWaves];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    const int n = S.nCellChunks;
    for (int b = 0; b < n; b += kBlock) {
        const int t = b + threadIdx.x;
        const int v = (t < n) ? ldg(S.cellChunkSum, (size_t)t) : 0;
        const int incl = wave_incl_scan(v);
        const int w = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 63) wsum[w] = incl;
        __syncthreads();
        int off = carry;
        for (int q = 0; q < w; ++q) off += wsum[q];
        if (t < n) G(S.cellChunkSum)[t] = off + incl - v;
        __syncthreads();
        if (threadIdx.x == kBlock - 1) carry = off + incl;
        __syncthreads();
    }
}

// Final scan: cellStart[c] = exclusive prefix of counts; counts reset.
// FUSED: the chunk totals are still the raw ones (no k_scan_offsets ran),
// and the block adds up those before its chunk itself; integer sums, so
// either way gives the same offsets.
template <bool FUSED>
__global__ __launch_bounds__(kBlock) void k_scan_final(const DevScene* __restrict__ T, const Chunk* __restrict__ C) {
    const Chunk c = C[blockIdx.x];
    const DevScene& S = T[c.scene];
    __shared__ int wsum[kWaves];
    __shared__ int base;
    const int chunk = c.begin / kCellChunk;
    if (FUSED) {
        int acc = 0;
        for (int t = threadIdx.x; t < chunk; t += kBlock) acc += ldg(S.cellChunkSum, (size_t)t);
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o);
        if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            int b = 0;
#pragma unroll
            for (int q = 0; q < kWaves; ++q) b += wsum[q];
            base = b;
        }
        __syncthreads();
    }
    const int c0 = c.begin + threadIdx.x * kCellPerThread;
    int vals[kCellPerThread];
    int local = 0;
#pragma unroll
    for (int u = 0; u < kCellPerThread; ++u) {
        const int cc = c0 + u;
        vals[u] = (cc < S.ncells) ? ldg(S.cellCount, (size_t)cc) : 0;
        local += vals[u];
    }
    const int incl = wave_incl_scan(local);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) wsum[w] = incl;
    __syncthreads();
    int woff = 0;
    for (int q = 0; q < w; ++q) woff += wsum[q];
    int run = (FUSED ? base : ldg(S.cellChunkSum, (size_t)chunk)) + woff + incl - local;
#pragma unroll
    for (int u = 0; u < kCellPerThread; ++u) {
        const int cc = c0 + u;
        if (cc < S.ncells) {
            G(S.cellStart)[cc] = run;
            G(S.cellCount)[cc] = 0;
        }
        run += vals[u];
    }
    // the chunk table covers ncells + 1 entries: the last one is the end
    if (c0 <= S.ncells && S.ncells < c0 + kCellPerThread) G(S.cellStart)[S.ncells] = S.n;
}

// Scatter the particle ids into their cells' slots (arrival order).
__global__ __launch_bounds__(kBlock) void k_scatter(const DevScene* __restrict__ T, const Chunk* __restrict__ C) {
    const Chunk c = C[blockIdx.x];
    const DevScene& S = T[c.scene];
    const int i = c.begin + threadIdx.x;
    const int cl = (i < S.n) ? ldg(S.cellU, (size_t)i) : -1;
    const int r = wave_count_add(S.cellCount, cl);
    if (i < S.n) G(S.perm)[ldg(S.cellStart, (size_t)cl) + r] = ldg(S.id, (size_t)i);
}

// Final slot of each particle = cell start + number of particles of its
// cell with a smaller id, so every bucket lists its particles in ascending
// id, the reference's insertion order (src/SPHSolver.cpp:188-191,
// src/HashTable.h:79-86).  The state is written straight to that slot (the
// other half of each double buffer); the cell counters are reset for the
// next step.
__global__ __launch_bounds__(kBlock) void k_rank_reorder(const DevScene* __restrict__ T, const Chunk* __restrict__ C) {
    const Chunk c = C[blockIdx.x];
    const DevScene& S = T[c.scene];
    const int i = c.begin + threadIdx.x;
    if (i >= S.n) return;
    const int cl = ldg(S.cellU, (size_t)i);
    const int s = ldg(S.cellStart, (size_t)cl), e = ldg(S.cellStart, (size_t)cl + 1);
    const uint32_t me = ldg(S.id, (size_t)i);
    // the particle's data is loaded before the rank loop and before any
    // store (a wait for a load also waits for the stores issued before it)
    const float4 xo = ld4(S.x, (size_t)i), vo = ld4(S.v, (size_t)i);
    const bool df = S.method == DFSPH;
    const float ko = df ? ldg(S.k, (size_t)i) : 0.f, kvo = df ? ldg(S.kv, (size_t)i) : 0.f;
    const float4 no = (df && S.pbOn) ? ld4(S.pbn, (size_t)i) : make_float4(0.f, 0.f, 0.f, 0.f);
    // rank within the cell: the ids of its first 8 members are loaded
    // together (a fluid cell holds ~8; the cell holds this particle, so
    // e > s and the clamped index stays in it), the rest one at a time
    uint32_t q8[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) q8[u] = ldg(S.perm, (size_t)min(s + u, e - 1));
    int rank = 0;
#pragma unroll
    for (int u = 0; u < 8; ++u) rank += ((s + u < e) & (q8[u] < me)) ? 1 : 0;
    for (int q = s + 8; q < e; ++q) rank += (ldg(S.perm, (size_t)q) < me) ? 1 : 0;
    const int d = s + rank;
    st4(S.x2, (size_t)d, xo);
    st4(S.v2, (size_t)d, vo);
    stg(S.id2, (size_t)d, me);
    stg(S.cell, (size_t)d, cl);
    if (df) {
        stg(S.k2, (size_t)d, ko);
        stg(S.kv2, (size_t)d, kvo);
        if (S.pbOn) st4(S.pbn2, (size_t)d, no);
    }
    if (S.pbOn) stg(S.everOcc, (size_t)cl, (uint8_t)1);
    stg(S.cellCount, (size_t)cl, 0);
}

// =========================================================== boundary

// Akinci boundary volume V_b = 1 / sum_b' W(x_b - x_b'), self included
// (src/AkinciBoundayModel.cpp:11-40).  One scene per launch; bvol sorted.
__global__ __launch_bounds__(kBlock) void k_bvolume(const DevScene* __restrict__ T, int scene, float* __restrict__ bvol) {
    const DevScene& S = T[scene];
    const int b = blockIdx.x * kBlock + threadIdx.x;
    if (b >= S.nb) return;
    const float4 xb = S.xb[b];
    const int cb = cell_of(S, xb);
    const KCoefD k = S.kd;
    const double bx = xb.x, by = xb.y, bz = xb.z;
    double sum = 0.0;
    walk27(S.bCellStart, cb, S.nx, S.nx * S.ny, [&](int j) {
        const Pair p = pair_of(bx, by, bz, S.xb[j], k);
        sum += W_d(p.q, k);
    });
    bvol[b] = (float)(1.0 / sum);
}
__global__ __launch_bounds__(kBlock) void k_bvolume_pack(const DevScene* __restrict__ T, int scene, const float* __restrict__ bvol) {
    const DevScene& S = T[scene];
    const int b = blockIdx.x * kBlock + threadIdx.x;
    if (b >= S.nb) return;
    S.xb[b].w = bvol[b];
}

// ================================================ neighbor list + density

// Per particle: walk the 27 cells, keep the fluid neighbors with
// |r|/h <= 1 decided exactly in FP64 (the reference's W and gradW vanish
// beyond h, so its sums over the whole bucket equal sums over this list),
// and in the same pass form
//   rho_i   = m W(0) + sum_j m W_ij + sum_b rho0 V_b W_ib   (src/SPHSolver.cpp:60-89)
//   alpha_i = rho_i / (|sum_j m gradW_ij + sum_b rho0 V_b gradW_ib|^2 + sum_j |m gradW_ij|^2)
//             or 0 when that is <= 1e-5                     (src/DFSPHSolver.cpp:139-186)
//   gb_i    = (sum_b rho0 V_b gradW_ib, sum_b rho0 V_b W_ib)
// plus, per method, the pressure of the Tait EOS (src/WCSPHSolver.cpp:93),
// the Spiky boundary sum used by PCISPH (src/PCISPHSolver.cpp:342-353) and
// the adhesion sum (src/AdhesionForce.h:46-59).
// Fluid neighbors of particle i in two passes.  Pass 1 walks the 27 cells
// with an FP32 distance test whose bound is 1e-5 above h^2 (FP32 rounding of
// |x_i - x_j|^2 is ~1e-7 relative) and writes the survivors to the list.
// Pass 2 runs the FP64 test and sums over those survivors only, compacting
// the list in place.  Pass 2 thus iterates ~1.1x the neighbor count on
// every lane of a wavefront, instead of running the FP64 path for every
// candidate that any lane of the wavefront accepts.  A particle with more
// survivors than list slots takes the one-pass walk over all candidates.
//
// Pass 1 reads its candidates from LDS.  The 64 particles of a wavefront are
// sorted by cell, so they span one cell range [cF, cL], and each of a lane's
// 9 candidate rows (cells c-1..c+1 of row (dy, dz)) lies inside that row's
// range cF-1..cL+1 for the whole wavefront.  The wavefront copies those 9
// particle ranges (~800 positions in a dense fluid, 12 B each) into LDS
// once, instead of every lane loading its ~216 candidates from memory.  A
// wavefront whose ranges exceed the stage reads global memory.
#ifndef SPHG_BUILD_STAGE
#define SPHG_BUILD_STAGE 1
#endif
constexpr int kWaveStage = SPHG_BUILD_STAGE ? 816 : 1;   // candidates per wavefront: 4 x 9.8 KB per block
// Pass 1 reads a staged row 4 candidates at a time from one LDS address (the
// 4 reads take immediate offsets); the last pass of a lane's row may read up
// to 3 entries past the staged ones, so the stage has 3 more.
constexpr int kWaveStagePad = kWaveStage + 3;

template <int M>
__global__ __launch_bounds__(kBlock) void k_build(const DevScene* __restrict__ T, const Chunk* __restrict__ C) {
    const Chunk c = C[blockIdx.x];
    const DevScene& S = T[c.scene];
    __shared__ float2 sxy[kBlock / 64][kWaveStagePad];
    __shared__ float sz[kBlock / 64][kWaveStagePad];
    const int i = c.begin + threadIdx.x;
    const bool valid = i < S.n;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const KCoefD k = S.kd;
    const float hh = S.kc.hhi2 * (1.0f + 1e-5f);   // conservative FP32 pre-filter
    const float4 xi = valid ? ld4(S.x, i) : make_float4(0.f, 0.f, 0.f, 0.f);
    const double xx = xi.x, xy = xi.y, xz = xi.z;
    const int ci = valid ? ldg(S.cell, (size_t)i) : 0;
    const int nx = S.nx, nxny = S.nx * S.ny;

    // the wavefront's 9 row ranges: first particle and offset in the stage
    const int wb = c.begin + (w << 6);
    const bool live = wb < S.n;
    int rfirst = 0, rcnt = 0;
    if (live && lane < 9) {
        const int cF = ldg(S.cell, (size_t)wb), cL = ldg(S.cell, (size_t)min(wb + 63, S.n - 1));
        const int off = (lane % 3 - 1) * nx + (lane / 3 - 1) * nxny;
        rfirst = ldg(S.cellStart, (size_t)(cF - 1 + off));
        rcnt = ldg(S.cellStart, (size_t)(cL + 2 + off)) - rfirst;
    }
    int incl = rcnt;
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
        const int t = __shfl_up(incl, o);
        if (lane >= o) incl += t;
    }
    int first[9], base[10];
    base[0] = 0;
#pragma unroll
    for (int r = 0; r < 9; ++r) {
        first[r] = __builtin_amdgcn_readlane(rfirst, r);
        base[r + 1] = __builtin_amdgcn_readlane(incl, r);
    }
    // the leading rows that fit are staged (all 9 in a dense fluid); the
    // rest of an oversized wavefront's rows are read from global memory
    int nStaged = 0;
    if (SPHG_BUILD_STAGE && live) {
#pragma unroll
        for (int r = 0; r < 9; ++r)
            if (base[r + 1] <= kWaveStage) nStaged = r + 1;
    }
    {
        // 4 loads per lane in flight before the first is written to LDS
        const int tot = base[nStaged];
        for (int t0 = lane; t0 < tot; t0 += 256) {
            float4 p[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int t = t0 + 64 * u;
                int f = first[0], b0 = 0;
#pragma unroll
                for (int q = 1; q < 9; ++q)
                    if (t >= base[q]) { f = first[q]; b0 = base[q]; }
                p[u] = (t < tot) ? ld4(S.x, (size_t)(f + (t - b0))) : make_float4(0.f, 0.f, 0.f, 0.f);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int t = t0 + 64 * u;
                if (t < tot) {
                    sxy[w][t] = make_float2(p[u].x, p[u].y);
                    sz[w][t] = p[u].z;
                }
            }
        }
    }
    __syncthreads();

    uint32_t* L = list_of_w(S, valid ? i : (i & ~63));
    const int maxN = S.maxNbr;
    // Pass 1 stores at byte offsets from the wavefront's group list (a
    // uniform base, so each store takes the base + 32-bit offset form; a
    // group's list is maxN KB).  The cursor is the offset of the lane's
    // next slot with the lane field (bits 4..9) held at all ones: + 4 then
    // carries from the last slot of a 4-slot block straight into the next
    // block (bits 10 up), and the store address is the cursor with the lane
    // field put back.  Its bits 2..3 and 10 up count the survivors.
    char* const Lw = (char*)(S.nbr + (size_t)__builtin_amdgcn_readfirstlane(i >> 6) * (size_t)maxN * kGroup);
    constexpr uint32_t kLaneBits = 0x3F0u;
    const uint32_t lane16 = (uint32_t)lane << 4;
    uint32_t cur = kLaneBits;
    const uint32_t curMax = kLaneBits | (slot_of(maxN - 1) << 2);
    // past the last slot the particle takes the one-pass walk, which
    // rewrites its list: the store is clamped to the last slot
    auto put = [&](uint32_t j) {
        const uint32_t a = (min(cur, curMax) & ~kLaneBits) | lane16;
        cur = (cur + 4u) | kLaneBits;
        *(gptr<uint32_t>)((gptr<char>)Lw + a) = j;
    };
    auto putSlot = [&](int s, uint32_t j) { *(gptr<uint32_t>)((gptr<char>)Lw + ((slot_of(s) << 2) | lane16)) = j; };
    auto curCount = [&]() { return (int)(((cur >> 10) << 2) | ((cur >> 2) & 3u)); };
    // PCISPH: the shell, candidates within h + skin that are not in the
    // list, fills the slots from the last one down; it stands in for the
    // 27-cell walk of the predicted densities while it fits beside the list
    float hsk = 0.f;
    if (M == PCISPH && S.pciShell) {
        const float hs = S.kc.h + pci_skin(S);
        hsk = hs * hs * (1.0f + 1e-4f);
    }
    int ns = 0;
    bool shellLost = false;
    int cnt = 0, ngr = 0;
    double ws = 0.0, sgx = 0.0, sgy = 0.0, sgz = 0.0, ssq = 0.0;
    // DFSPH: sum_j g_ij (v_i - v_j).(x_i - x_j) of the divergence error
    double dv = 0.0;
    const float4 vi = (M == DFSPH && valid) ? ld4(S.v, i) : make_float4(0.f, 0.f, 0.f, 0.f);
    const double ux = vi.x, uy = vi.y, uz = vi.z;
    // FP64 test and sums of one candidate that passed the FP32 test
    int nc = 0;   // FP32 survivors of pass 1 (slots [0, nc) while pass 2 reads them)
    // a survivor the FP64 test rejects goes to the shell (PCISPH)
    auto toShell = [&](uint32_t j) {
        if (M != PCISPH) return;
        if (ns < maxN - nc) stg(L, (size_t)slot_of(maxN - 1 - ns), j);
        else shellLost = true;
        ++ns;
    };
    // pos: the slot the candidate was read from (-1: from the walk); an
    // entry is written only where the compaction moves it
    auto visit = [&](uint32_t j, const float4& xj, const float4& vj, int pos) {
        const Pair p = pair_of(xx, xy, xz, xj, k);
        if (!(p.r2 <= k.hhi2)) {
            toShell(j);
            return;
        }
        // 1e-9 <= r2 < hlo2: |r| > 1e-5 and q < 1, where W and the gradient
        // take their inner forms and the gradient is nonzero.  The
        // reference's gradient (src/CubicSpline.h:55-73) vanishes only at
        // |r| <= 1e-5 and q = 1; near those two, and near q = 1 for the list
        // membership, the test is repeated operation-exact in FP64.
        double g = G_in(p, k), wq = W_in(p.q, k);
        bool gnz = true;
        if (p.r2 >= k.hlo2 || p.r2 < 1e-9) {
            const Exact e = exact_q(xi, xj, S.hD);
            if (!(e.q <= 1.0)) {
                toShell(j);
                return;
            }
            gnz = grad_nonzero64(e);
            g = G_d(p, k);
            wq = W_d(p.q, k);
        }
        if (cnt < maxN && cnt != pos) stg(L, (size_t)slot_of(cnt), j);
        ++cnt;
        ngr += gnz ? 1 : 0;
        ws += wq;
        sgx += g * p.dx;
        sgy += g * p.dy;
        sgz += g * p.dz;
        ssq += g * g * p.r2;
        if (M == DFSPH) dv += g * ((ux - (double)vj.x) * p.dx + (uy - (double)vj.y) * p.dy + (uz - (double)vj.z) * p.dz);
    };
    auto vel = [&](uint32_t j) { return (M == DFSPH) ? ld3o(S.v, j) : make_float4(0.f, 0.f, 0.f, 0.f); };
    // pass 1 over the 27 cells as 9 rows (walk27 order), candidate positions
    // from `fetch(j, row)`
    // (4 candidates per pass: their loads are issued before the first test)
    auto scanRow = [&](int r, int s0, int e0, int& nc, auto&& fetch) {
        for (int j0 = s0; j0 < e0; j0 += 4) {
            float3 q[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) q[u] = fetch(min(j0 + u, e0 - 1), r);
            bool ok[4], sh[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int j = j0 + u;
                const float fx = xi.x - q[u].x, fy = xi.y - q[u].y, fz = xi.z - q[u].z;
                const float d2 = fx * fx + fy * fy + fz * fz;
                // evaluated without short-circuit: one exec-mask section
                ok[u] = (j < e0) & (j != i) & (d2 <= hh);
                sh[u] = (M == PCISPH) & (j < e0) & (j != i) & (d2 > hh) & (d2 <= hsk);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (ok[u]) {
                    put((uint32_t)(j0 + u));
                    if (M == PCISPH) ++nc;
                }
                if (M == PCISPH && sh[u]) {
                    // below the list's slots only (list writes that later
                    // reach the shell are caught by nc + ns > maxN)
                    if (nc + ns < maxN) putSlot(maxN - 1 - ns, (uint32_t)(j0 + u));
                    else shellLost = true;
                    ++ns;
                }
            }
        }
    };
    // pass 1 over a staged row: candidate j sits at stage entry
    // j + base[r] - first[r], and one LDS address per row advances by 4
    // entries a pass, so the candidates need no per-read index arithmetic.
    // Entries past the lane's row end (at most 3, inside the padded stage)
    // are read and fail the u < e0 - j0 test.  Particle i itself can only
    // lie in row 4 (cells ci-1..ci+1): the other rows are offset by at
    // least nx >= 3 cells, so only that row tests j != i.
    auto scanRowLds = [&](auto rc, int s0, int e0, int& nc) {
        constexpr int r = decltype(rc)::value;
        const int t0 = s0 + (base[r] - first[r]);
        const float2* pxy = &sxy[w][t0];
        const float* pz = &sz[w][t0];
        for (int j0 = s0; j0 < e0; j0 += 4, pxy += 4, pz += 4) {
            float d2[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const float2 a = pxy[u];
                const float fx = xi.x - a.x, fy = xi.y - a.y, fz = xi.z - pz[u];
                d2[u] = fx * fx + fy * fy + fz * fz;
            }
            const int rem = e0 - j0;
            bool ok[4], sh[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const bool in = (u == 0 || u < rem) & (r != 4 || j0 + u != i);
                ok[u] = in & (d2[u] <= hh);
                sh[u] = (M == PCISPH) & in & (d2[u] > hh) & (d2[u] <= hsk);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (ok[u]) {
                    put((uint32_t)(j0 + u));
                    if (M == PCISPH) ++nc;
                }
                if (M == PCISPH && sh[u]) {
                    if (nc + ns < maxN) putSlot(maxN - 1 - ns, (uint32_t)(j0 + u));
                    else shellLost = true;
                    ++ns;
                }
            }
        }
    };
    auto gblFetch = [&](int j, int) {
        const float4 p = ld4(S.x, (size_t)j);
        return make_float3(p.x, p.y, p.z);
    };
    if (valid) {
        // the 9 rows' ranges are loaded before pass 1's first list store:
        // a wait for a load also waits for every store issued before it
        int rs[9], re[9];
#pragma unroll
        for (int r = 0; r < 9; ++r) {
            const int off = (r % 3 - 1) * nx + (r / 3 - 1) * nxny;
            rs[r] = ldg(S.cellStart, (size_t)(ci - 1 + off));
            re[r] = ldg(S.cellStart, (size_t)(ci + 2 + off));
        }
        static_for<9>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            if (r < nStaged) scanRowLds(rc, rs[r], re[r], nc);
            else scanRow(r, rs[r], re[r], nc, gblFetch);
        });
        nc = curCount();
        if (M == PCISPH && nc + ns > maxN) shellLost = true;
        if (nc <= maxN) {
            // 4 survivors per pass, their list entries and positions loaded
            // first; visit writes slots <= the one it reads, so entries
            // already loaded are not overwritten before use, and the next
            // block of entries is requested before this pass's data is used
            v4u bn = ld_block(L, 0);
            for (int q0 = 0; q0 < nc; q0 += 4) {
                // one block of 4 slots; those past the last survivor hold
                // stale values and are replaced by the first (a valid index)
                const v4u blk = bn;
                const uint32_t jj[4] = {blk.x, q0 + 1 < nc ? blk.y : blk.x, q0 + 2 < nc ? blk.z : blk.x,
                                        q0 + 3 < nc ? blk.w : blk.x};
                float4 xj[4], vj[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    xj[u] = ld3o(S.x, jj[u]);
                    vj[u] = vel(jj[u]);
                }
                bn = ld_block(L, (q0 + 4 < nc ? q0 + 4 : q0) >> 2);
                asm volatile("" ::: "memory");
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (q0 + u < nc) visit(jj[u], xj[u], vj[u], q0 + u);
            }
        } else {
            shellLost = true;
            ns = maxN;   // toShell writes nothing from here on
            walk27(S.cellStart, ci, nx, nxny, [&](int j) {
                const float4 xj = ld4(S.x, j);
                const float fx = xi.x - xj.x, fy = xi.y - xj.y, fz = xi.z - xj.z;
                if (!(fx * fx + fy * fy + fz * fz <= hh) || j == i) return;
                visit((uint32_t)j, xj, vel((uint32_t)j), -1);
            });
        }
    }
    // pad the list with the particle itself up to the group's longest list
    // rounded to 4 (a self term is exactly zero in every sweep)
    const int own = min(cnt, maxN);
    int pad = min((own + 3) & ~3, maxN);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) pad = max(pad, __shfl_xor(pad, o));
    double e = 0.0;   // DFSPH: this particle's share of the average divergence error
    if (valid) {
    double bw = 0.0, bgx = 0.0, bgy = 0.0, bgz = 0.0;
    double sx2 = 0.0, sy2 = 0.0, sz2 = 0.0;   // spiky boundary sum (PCISPH)
    double ax = 0.0, ay = 0.0, az = 0.0;      // adhesion sum
    const bool wantAdh = (S.flags & F_ADH) != 0;
    const double rho0 = S.rho0D;
    // boundary sums, in two passes like the fluid: FP32 survivors go to the
    // free list slots past the fluid list, then the FP64 pass runs over them
    auto bvisit = [&](const float4& y) {
        const Pair p = pair_of(xx, xy, xz, y, k);
        if (!(p.q <= 1.0)) return;
        const double cb = rho0 * (double)y.w;
        bw += cb * W_d(p.q, k);
        const double g = cb * G_d(p, k);
        bgx += g * p.dx;
        bgy += g * p.dy;
        bgz += g * p.dz;
        if (M == PCISPH) {
            const double gs = cb * Gs_d(p, k);
            sx2 += gs * p.dx;
            sy2 += gs * p.dy;
            sz2 += gs * p.dz;
        }
        if (wantAdh && p.r2 > 0.0) {
            const double sa = cb * A_d(p.rl, k) * p.inv;
            ax += sa * p.dx;
            ay += sa * p.dy;
            az += sa * p.dz;
        }
    };
    // the boundary survivors use the free slots between the list and the shell
    const int room = max(0, maxN - own - ((M == PCISPH && !shellLost) ? ns : 0));
    int nbc = 0;
    {
        // the 9 rows' ranges first, then 4 candidates' loads at a time: a
        // wait for a load also waits for the survivor stores before it
        int bs[9], be[9];
#pragma unroll
        for (int r = 0; r < 9; ++r) {
            const int off = (r % 3 - 1) * nx + (r / 3 - 1) * nxny;
            bs[r] = ldg(S.bCellStart, (size_t)(ci - 1 + off));
            be[r] = ldg(S.bCellStart, (size_t)(ci + 2 + off));
        }
#pragma unroll 1
        for (int r = 0; r < 9; ++r) {
            for (int b0 = bs[r]; b0 < be[r]; b0 += 4) {
                float4 y[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) y[u] = ld4(S.xb, min(b0 + u, be[r] - 1));
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const float fx = xi.x - y[u].x, fy = xi.y - y[u].y, fz = xi.z - y[u].z;
                    if ((b0 + u < be[r]) & (fx * fx + fy * fy + fz * fz <= hh)) {
                        if (nbc < room) putSlot(own + nbc, (uint32_t)(b0 + u));
                        ++nbc;
                    }
                }
            }
        }
    }
    if (nbc <= room) {
        for (int q0 = 0; q0 < nbc; q0 += 4) {
            float4 yb[4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                yb[u] = ld4(S.xb, ldg((const uint32_t*)L, (size_t)slot_of(own + min(q0 + u, nbc - 1))));
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (q0 + u < nbc) bvisit(yb[u]);
        }
    } else {
        walk27(S.bCellStart, ci, nx, nxny, [&](int b) {
            const float4 y = ld4(S.xb, b);
            const float fx = xi.x - y.x, fy = xi.y - y.y, fz = xi.z - y.z;
            if (fx * fx + fy * fy + fz * fz <= hh) bvisit(y);
        });
    }
    // pad the list with the particle itself up to the group's longest list
    // (overwrites the boundary survivors' slots)
    for (int q = own; q < pad; ++q) stg(L, (size_t)slot_of(q), (uint32_t)i);
    if (M == PCISPH) {
        if (pad + ns > maxN) shellLost = true;   // the padding reached the shell
        stg(S.nsh, (size_t)i, (uint16_t)(shellLost ? 0xFFFF : ns));
    }
    if ((i & 63) == 0) stg(S.gcnt, (size_t)(i >> 6), (uint16_t)pad);
    if (cnt > maxN) {
        atomicMax(&S.ss->overflow, (unsigned)cnt);
        atomicMax(&S.ss->listOver, (unsigned)cnt);
    }
    stg(S.nnb, (size_t)i, (uint16_t)own);
    stg(S.ngrad, (size_t)i, (uint8_t)min(ngr, 255));
    const double m = S.mD;
    const double drho = (m * (k.ck + ws) + bw) - rho0;
    const double rho = rho0 + drho;
    stg(S.drho, (size_t)i, (float)drho);
    st4(S.gb, i, make_float4((float)bgx, (float)bgy, (float)bgz, (float)bw));
    if (wantAdh) st4(S.adh, i, make_float4((float)ax, (float)ay, (float)az, 0.f));
    if (M == DFSPH) {
        const double tx = m * sgx + bgx, ty = m * sgy + bgy, tz = m * sgz + bgz;
        const double den = tx * tx + ty * ty + tz * tz + m * m * ssq;
        const float alpha = (fabs(den) > 1e-5) ? (float)(rho / den) : 0.0f;
        stg(S.alpha, (size_t)i, alpha);
        // initial divergence error of the step's divergence solve
        // (computeDivergenceError, src/DFSPHSolver.cpp:408-492, and kappa_v,
        // :531), from the pairs above; stored values are read back as the
        // later sweeps read them (FP32 gb, delta rho, alpha)
        const float4 gbf = make_float4((float)bgx, (float)bgy, (float)bgz, (float)bw);
        const double dc = m * dv + (ux * (double)gbf.x + uy * (double)gbf.y + uz * (double)gbf.z);
        const double de = (ngr < 20) ? 0.0 : fmax(dc, 0.0);
        const double rhof = rho0 + (double)(float)drho;
        // the divergence of the velocities: the next step's D0 unless a
        // divergence iteration replaces it
        ((gptr<double>)S.d0)[i] = dc;
        stg(S.divErr, (size_t)i, (float)de);
        st4(S.xk, i, make_float4(xi.x, xi.y, xi.z, (float)(de * (double)alpha / (S.ss->dt * rhof))));
        e = de / rho0;
    } else if (M == WCSPH) {
        // p = max(0, kappa rho0/gamma ((rho/rho0)^gamma - 1)), written
        // (1+x)^g - 1 = expm1(g log1p(x)) with x = (rho - rho0)/rho0
        const double x = drho / rho0;
        const double pr = fmax(0.0, (double)S.stiffness * rho0 / (double)S.gammaEos * expm1((double)S.gammaEos * log1p(x)));
        stg(S.p, (size_t)i, (float)pr);
        st4(S.xk, i, make_float4(xi.x, xi.y, xi.z, (float)(pr / (rho * rho))));
        const float4 vi = ld4(S.v, i);
        st4(S.aux, i, make_float4(vi.x, vi.y, vi.z, (float)(m / rho)));
    } else {
        st4(S.gbs, i, make_float4((float)sx2, (float)sy2, (float)sz2, 0.f));
    }
    }
    if (M == DFSPH) {
        __shared__ double sh[kWaves];
        const double bs = block_sum(e, sh);
        if (threadIdx.x == 0) S.partial[c.begin / kBlock] = bs;
        if (c.begin == 0 && threadIdx.x == 0) S.ss->d0Valid = 1u;
    }
}

// ======================================================= non-pressure

// Akinci surface normals n_i = h sum_j (m/rho_j) gradW_ij
// (src/SurfaceTensionAkinci.h:46-61).  nrm.w carries rho_i - rho0 for pass 2.
__global__ __launch_bounds__(kBlock) void k_normals(const DevScene* __restrict__ T, const Chunk* __restrict__ C) {
    const Chunk c = C[blockIdx.x];
    const DevScene& S = T[c.scene];
    const int i = c.begin + threadIdx.x;
    if (i >= S.n) return;
    const KCoefD k = S.kd;
    const float4 xi = ld4(S.x, i);
    const double xx = xi.x, xy = xi.y, xz = xi.z;
    const double m = S.mD, rho0 = S.rho0D;
    double nx = 0.0, ny = 0.0, nz = 0.0;
    sweep4(S, i, [&](uint32_t j) { XV r; r.x = ld4(S.x, j); r.v.x = ldg(S.drho, j); return r; },
           [&](uint32_t, const XV& d) {
               const Pair p = pair_of(xx, xy, xz, d.x, k);
               const double g = G_d(p, k) * (m / (rho0 + (double)d.v.x));
               nx += g * p.dx;
               ny += g * p.dy;
               nz += g * p.dz;
           });
    const double h = k.h;
    st4(S.nrm, i, make_float4((float)(nx * h), (float)(ny * h), (float)(nz * h), ldg(S.drho, i)));
}

// The end of a WCSPH or PCISPH step: time += dt (src/WCSPHSolver.cpp:77,
// src/PCISPHSolver.cpp:80), and which predicted-density path a PCISPH step
// took.
__device__ inline void advance_time(const DevScene& S) {
    Scalars* ss = S.ss;
    ss->time += ss->dt;
    if (S.method == PCISPH) {
        if (S.pciShell && ss->shellBad == 0u) ss->shellSteps += 1u;
        else ss->walkSteps += 1u;
    }
}
// Whether the step's last kernel (the WCSPH forces, the PCISPH integration)
// advances the time, one launch fewer per step: not when the cube clamp, the
// PCISPH boundary projection or an emitter runs after it (the emitters read
// the step's start time); k_time_advance does it then.
__device__ inline bool time_in_last_kernel(const DevScene& S) { return !(S.cubeOn || S.pbOn || S.nEmit > 0); }

// Non-pressure accelerations (and, for WCSPH, the pressure force and the
// explicit integration).
//   pressure  src/WCSPHSolver.cpp:152-186
//   XSPH      src/XSPHViscosity.h:45-67
//   standard viscosity   src/StandardViscosityForce.h:37-74:
//     a_i += nu/rho0 sum_j m (v_j - v_i)/rho_j lapl(r_ij)
//          + nu_b/rho0 sum_b rho0 V_b (-v_i)/rho_i lapl(r_ib),
//     lapl(r) = 45/(pi h^6) (h - r) (src/ViscoK.h:28-35)
//   artificial viscosity src/ArtificialViscosity.h:40-81:
//     a_i += sum_j k_ij nu (m/rho_j) (v_i - v_j).r_ij/(r^2 + 0.01 h^2) gradW_ij
//          + sum_b nu_b (rho0 V_b/rho_i) v_i.r_ib/(r^2 + 0.01 h^2) gradW_ib,
//     k_ij = 2 rho0/(rho_i + rho_j); the reference's factor 10 * 0.1 is 1
//   (boundary terms only when nu_b != 0, walked here over the boundary cells)
//   cohesion  src/SurfaceTensionAkinci.h:66-96
//   adhesion  src/AdhesionForce.h:38-61
//   integrate src/SPHSolver.cpp:154-167
// M = WCSPH reads neighbor positions from xk and velocities from aux (both
// written by k_build) because it updates x and v in place.  M = PCISPH also
// predicts v*, x* for the first pressure iteration (pacc = 0,
// src/PCISPHSolver.cpp:139-146).
template <int M>
__global__ __launch_bounds__(kBlock) void k_forces(const DevScene* __restrict__ T, const Chunk* __restrict__ C) {
    const Chunk c = C[blockIdx.x];
    const DevScene& S = T[c.scene];
    const int i = c.begin + threadIdx.x;
    if ((M == WCSPH) && c.begin == 0 && threadIdx.x == 0 && time_in_last_kernel(S)) advance_time(S);
    if (i >= S.n) return;
    const KCoefD k = S.kd;
    const double dt = S.ss->dt;
    const bool xs = (S.flags & F_XSPH) != 0, st = (S.flags & F_ST) != 0;
    const bool vstd = (S.flags & F_VSTD) != 0, vart = (S.flags & F_VART) != 0;
    const bool vel = xs || vstd || vart;
    const float4 xi = (M == WCSPH) ? ld4(S.xk, i) : ld4(S.x, i);
    const float4 vi = (M == WCSPH) ? ld4(S.aux, i) : ld4(S.v, i);
    const double xx = xi.x, xy = xi.y, xz = xi.z;
    const double vx0 = vi.x, vy0 = vi.y, vz0 = vi.z;
    const double m = S.mD, rho0 = S.rho0D;
    const double rhoi = rho0 + (double)ldg(S.drho, i);
    const float4 ni = st ? ld4(S.nrm, i) : make_float4(0.f, 0.f, 0.f, 0.f);
    const double pri = xi.w;
    double px = 0.0, py = 0.0, pz = 0.0;   // pressure part
    double vx = 0.0, vy = 0.0, vz = 0.0;   // XSPH part (before nu/dt)
    double sx = 0.0, sy = 0.0, sz = 0.0;   // surface tension part (before gamma)
    double wx = 0.0, wy = 0.0, wz = 0.0;   // standard / artificial viscosity part
    sweep4(
        S, i,
        [&](uint32_t j) {
            NbF r;
            r.x = (M == WCSPH) ? ld4(S.xk, j) : ld4(S.x, j);
            if (vel) r.v = (M == WCSPH) ? ld4(S.aux, j) : ld4(S.v, j);
            // rho_j: XSPH under WCSPH takes m/rho_j from aux.w instead
            if ((vel && M != WCSPH) || vstd || vart) r.d = ldg(S.drho, j);
            if (st) r.n = ld4(S.nrm, j);
            return r;
        },
        [&](uint32_t, const NbF& d) {
            const Pair p = pair_of(xx, xy, xz, d.x, k);
            if (M == WCSPH) {
                const double g = G_d(p, k) * m * (pri + (double)d.x.w);
                px -= g * p.dx;
                py -= g * p.dy;
                pz -= g * p.dz;
            }
            if (xs) {
                const double mj = (M == WCSPH) ? (double)d.v.w : m / (rho0 + (double)d.d);
                const double w = mj * W_d(p.q, k);
                vx += w * ((double)d.v.x - vx0);
                vy += w * ((double)d.v.y - vy0);
                vz += w * ((double)d.v.z - vz0);
            }
            if (vstd) {
                // m/rho_j lapl(r) (v_j - v_i); list members have r <= h
                const double w = (m / (rho0 + (double)d.d)) * (k.vl * (k.h - p.rl));
                wx += w * ((double)d.v.x - vx0);
                wy += w * ((double)d.v.y - vy0);
                wz += w * ((double)d.v.z - vz0);
            }
            if (vart) {
                const double rhoj = rho0 + (double)d.d;
                const double kij = 2.0 * rho0 / (rhoi + rhoj);
                const double dot = (vx0 - (double)d.v.x) * p.dx + (vy0 - (double)d.v.y) * p.dy + (vz0 - (double)d.v.z) * p.dz;
                const double f = kij * (m / rhoj) * dot / (p.rl * p.rl + 0.01 * k.h2) * G_d(p, k);
                wx += f * p.dx;
                wy += f * p.dy;
                wz += f * p.dz;
            }
            if (st && p.r2 > 0.0) {
                // list members satisfy |r| <= h exactly (FP64), the reference's
                // test at src/SurfaceTensionAkinci.h:88
                const double kij = 2.0 * rho0 / (rhoi + (rho0 + (double)d.n.w));
                bool outer = 2.0 * p.rl > k.h;
                if (fabs(2.0 * p.rl - k.h) < 1e-9 * k.h) outer = 2.0 * exact_q(xi, d.x, S.hD).rl > S.hD;
                const double cs = -m * C_d(p.rl, outer, k) * p.inv;
                sx += kij * (cs * p.dx - ((double)ni.x - (double)d.n.x));
                sy += kij * (cs * p.dy - ((double)ni.y - (double)d.n.y));
                sz += kij * (cs * p.dz - ((double)ni.z - (double)d.n.z));
            }
        });
    double ax = 0.0, ay = 0.0, az = 0.0;
    if (M == WCSPH) {
        const float4 gb = S.gb[i];
        ax = px - pri * (double)gb.x;
        ay = py - pri * (double)gb.y;
        az = pz - pri * (double)gb.z;
        if (S.pbOn) {
            // the boundary particles' own pressure term m_b p_b / rho_b^2
            // gradW of the PCISPH method (src/WCSPHSolver.cpp:160-172)
            walk27(S.bCellStart, ldg(S.cell, (size_t)i), S.nx, S.nx * S.ny, [&](int b) {
                const float4 y = ld4(S.xb, b);
                const Pair p = pair_of(xx, xy, xz, y, k);
                const double g = S.rho0D * (double)y.w * (double)S.bpr[b] * G_d(p, k);
                ax -= g * p.dx;
                ay -= g * p.dy;
                az -= g * p.dz;
            });
        }
    }
    if (xs) {
        const double f = (double)S.nu / dt;
        ax += f * vx;
        ay += f * vy;
        az += f * vz;
        if (S.nub != 0.0f) {
            const double fb = (double)S.nub * ((double)S.gb[i].w / rhoi) / dt;
            ax -= fb * vx0;
            ay -= fb * vy0;
            az -= fb * vz0;
        }
    }
    if (vstd || vart) {
        const double nu = S.nu, nub = S.nub;
        double bx = 0.0, by = 0.0, bz = 0.0;
        if (nub != 0.0) {
            walk27(S.bCellStart, ldg(S.cell, (size_t)i), S.nx, S.nx * S.ny, [&](int b) {
                const float4 y = ld4(S.xb, b);
                const Pair p = pair_of(xx, xy, xz, y, k);
                if (!(p.q <= 1.0)) return;
                const double cb = rho0 * (double)y.w;
                if (vstd) {
                    const double w = cb * (k.vl * (k.h - p.rl)) / rhoi;
                    bx -= w * vx0;
                    by -= w * vy0;
                    bz -= w * vz0;
                } else {
                    const double dot = vx0 * p.dx + vy0 * p.dy + vz0 * p.dz;
                    const double f = (cb / rhoi) * dot / (p.rl * p.rl + 0.01 * k.h2) * G_d(p, k);
                    bx += f * p.dx;
                    by += f * p.dy;
                    bz += f * p.dz;
                }
            });
        }
        const double fv = vstd ? nu / rho0 : nu, fb = vstd ? nub / rho0 : nub;
        ax += fv * wx + fb * bx;
        ay += fv * wy + fb * by;
        az += fv * wz + fb * bz;
    }
    if (st) {
        ax += (double)S.stCoef * sx;
        ay += (double)S.stCoef * sy;
        az += (double)S.stCoef * sz;
    }
    if (S.flags & F_ADH) {
        const float4 ad = S.adh[i];
        ax -= (double)S.adhBeta * (double)ad.x;
        ay -= (double)S.adhBeta * (double)ad.y;
        az -= (double)S.adhBeta * (double)ad.z;
    }
    S.a[i] = make_float4((float)ax, (float)ay, (float)az, 0.f);
    if (M == WCSPH) {
        const double nvx = vx0 + (ax + (double)S.gx) * dt;
        const double nvy = vy0 + (ay + (double)S.gy) * dt;
        const double nvz = vz0 + (az + (double)S.gz) * dt;
        S.v[i] = make_float4((float)nvx, (float)nvy, (float)nvz, 0.f);
        S.x[i] = make_float4((float)(xx + nvx * dt), (float)(xy + nvy * dt), (float)(xz + nvz * dt), 0.f);
    } else if (M == PCISPH) {
        const double wx = vx0 + (ax + (double)S.gx) * dt;
        const double wy = vy0 + (ay + (double)S.gy) * dt;
        const double wz = vz0 + (az + (double)S.gz) * dt;
        S.xs[i] = make_float4((float)(wx * dt), (float)(wy * dt), (float)(wz * dt), 0.f);
        if (S.pbOn) S.pv[i] = make_float4((float)wx, (float)wy, (float)wz, 0.f);
        shell_guard(S, wx * dt, wy * dt, wz * dt);
    }
}

// WCSPH/PCISPH step-time bookkeeping: steps++ (one thread per scene) and,
// for PCISPH, the entry test of the pressure loop (src/PCISPHSolver.cpp:365)
// with p = pacc = 0 (:83-120).

// ================================================================ DFSPH

// CFL time step (src/DFSPHSolver.cpp:827-870): maxVel = max(0.1, max_i
// sqrt(|v_i + (a_i + g) dt|)); dt' = cfl if cfl < dt, else 0.99 dt after
// more than 3 density iterations, else 1.001 dt; clamped to [min, max].
__global__ __launch_bounds__(kBlock) void k_dt(const DevScene* __restrict__ T, const Chunk* __restrict__ C) {
    const Chunk c = C[blockIdx.x];
    const DevScene& S = T[c.scene];
    const gptr<const Scalars> ss = G((const Scalars*)S.ss);
    __shared__ double sh[kWaves];
    const int i = c.begin + threadIdx.x;
    const double dt = ss->dt;
    double val = 0.0;
    if (i < S.n) {
        const float4 v = ld4(S.v, (size_t)i);
        double ax = S.gx, ay = S.gy, az = S.gz;
        if (S.flags & F_FORCES) {
            const float4 a = ld4(S.a, (size_t)i);
            ax = (double)a.x + ax;
            ay = (double)a.y + ay;
            az = (double)a.z + az;
        }
        const double cx = v.x + ax * dt, cy = v.y + ay * dt, cz = v.z + az * dt;
        val = sqrt(sqrt(cx * cx + cy * cy + cz * cz));
    }
    const double bm = block_max(val, sh);
    if (threadIdx.x == 0) G(S.partial)[c.begin / kBlock] = bm;
}

// The density solve (predict, warm start, density iterations) keeps the
// velocity of the step's start, v0, unchanged in v and accumulates the
// step's change u = v - v0 in x.w and vyz, the slots its error sweeps
// gather; the position update forms v = v0 + u.  The predicted density is
// rho + dt (D0 + U): D0, the sum over v0, is formed once per step in FP64,
// and U, the sum over u, in every iteration.  Velocities rounded to FP32
// after every sweep would put their rounding (~1e-7 of |v|) into rho*, which
// kappa = (rho* - rho0) alpha / (dt^2 rho) amplifies; u is the size of the
// step's corrections, so its FP32 rounding is far below that.
__device__ inline void store_u(const DevScene& S, int i, double ux, double uy, double uz) {
    ((gptr<float>)S.x)[4 * (size_t)i + 3] = (float)ux;
    ((gptr<v2f>)S.vyz)[i] = v2f{(float)uy, (float)uz};
}

// predictVelocities (src/DFSPHSolver.cpp:191-214) fused with the first loop
// of densityWarmStart (:602-613): k_i <- min(k_i/rho_i, 1e-9)/dt^2, staged
// in xk.w for the warm-start sweep.  The predicted velocity is kept as
// u = (a + g) dt beside v0 (and, with the PCISPH boundary method, the change
// of its velocity correction).
__global__ __launch_bounds__(kBlock) void k_df_predict(const DevScene* __restrict__ T, const Chunk* __restrict__ C) {
    const Chunk c = C[blockIdx.x];
    const DevScene& S = T[c.scene];
    const int i = c.begin + threadIdx.x;
    if (i >= S.n) return;
    const double dt = G(S.ss)->dt;
    const float4 v = ld4(S.v, (size_t)i);
    double ax = S.gx, ay = S.gy, az = S.gz;
    if (S.flags & F_FORCES) {
        const float4 a = ld4(S.a, (size_t)i);
        ax = (double)a.x + ax;
        ay = (double)a.y + ay;
        az = (double)a.z + az;
    }
    double ux = ax * dt, uy = ay * dt, uz = az * dt;
    if (S.pbOn) {
        // correctVelocities with the normals of the last step's projection
        // (src/DFSPHSolver.cpp:51-58)
        const float4 n = ld4(S.pbn, (size_t)i);
        double wx = (double)v.x + ux, wy = (double)v.y + uy, wz = (double)v.z + uz;
        if (pb_velocity(S, n.x, n.y, n.z, wx, wy, wz)) {
            ux = wx - (double)v.x;
            uy = wy - (double)v.y;
            uz = wz - (double)v.z;
        }
    }
    const double rho = S.rho0D + (double)ldg(S.drho, (size_t)i);
    const double kw = -fmax(-(double)ldg(S.k, (size_t)i) / rho, -1.0e-9) / (dt * dt);
    const float4 x = ld4(S.x, (size_t)i);
    st4(S.xk, (size_t)i, make_float4(x.x, x.y, x.z, (float)kw));
    store_u(S, i, ux, uy, uz);
}

// Velocity correction sweep shared by the warm start (src/DFSPHSolver.cpp:
// 622-677), the density solve (:335-395) and the divergence solve
// (:520-579):  v_i -= dt (sum_j m (s_i + s_j) gradW_ij [|s_i+s_j| > eps]
//                         + c s_i gb_i [|c s_i| > eps])
// with s = xk.w, c = 2 for the warm start, 1 otherwise.  MODE 0 = warm start
// (then k_i = 0, :679-689), 1 = density (k_i += (rho*_i - rho0) alpha_i,
// :349), 2 = divergence (kv_i += divErr_i alpha_i, :534).
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_df_vel(const DevScene* __restrict__ T, const Chunk* __restrict__ C) {
    const Chunk c = C[blockIdx.x];
    const DevScene& S = T[c.scene];
    const Scalars* ss = S.ss;
    if (MODE == 1 && ss->done) return;
    if (MODE == 2 && ss->doneV) return;
    const int i = c.begin + threadIdx.x;
    if (i >= S.n) return;
    const KCoefD k = S.kd;
    const double eps = 1.0e-5;
    const double dt = ss->dt;
    const float4 xi = ld4(S.xk, i);
    const double xx = xi.x, xy = xi.y, xz = xi.z;
    const double si = xi.w;
    const double m = S.mD;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    sweep4(S, i, [&](uint32_t j) { return ld4o(S.xk, j); },
           [&](uint32_t, const float4& xj) {
               const double ks = si + (double)xj.w;
               const PairG p = pairg_of(xx, xy, xz, xj, k);
               const double g = sel(fabs(ks) > eps, m * ks * p.g, 0.0);
               sx += g * p.dx;
               sy += g * p.dy;
               sz += g * p.dz;
           });
    // every load of the epilogue is issued before its first store: a wait
    // for a load also waits for the stores issued before it
    const float4 gb = S.gb[i];
    float kold = 0.f, kerr = 0.f, kal = 0.f;
    if (MODE == 1) { kold = S.k[i]; kerr = S.drhoS[i]; kal = S.alpha[i]; }
    if (MODE == 2) { kold = S.kv[i]; kerr = S.divErr[i]; kal = S.alpha[i]; }
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    float uxf = 0.f;
    float2 uyz = make_float2(0.f, 0.f);
    if (MODE <= 1) {
        uxf = ((gptr<const float>)S.x)[4 * (size_t)i + 3];
        uyz = ld2o(S.vyz, (uint32_t)i);
    } else {
        v = S.v[i];
    }
    const double cb = (MODE == 0) ? 2.0 * si : si;
    if (fabs(cb) > eps) {
        sx += cb * (double)gb.x;
        sy += cb * (double)gb.y;
        sz += cb * (double)gb.z;
    }
    if (MODE <= 1) {
        // the density solve: the change u (x.w, vyz; no kernel gathers x
        // during this sweep) takes the correction, v keeps v0
        store_u(S, i, (double)uxf - dt * sx, (double)uyz.x - dt * sy, (double)uyz.y - dt * sz);
    } else {
        v.x = (float)((double)v.x - dt * sx);
        v.y = (float)((double)v.y - dt * sy);
        v.z = (float)((double)v.z - dt * sz);
        S.v[i] = v;
        // copies for the error sweep's gathers: v_x into the unused x.w (no
        // kernel gathers x during this sweep), (v_y, v_z) into vyz
        ((gptr<float>)S.x)[4 * (size_t)i + 3] = v.x;
        ((gptr<v2f>)S.vyz)[i] = v2f{v.y, v.z};
    }
    if (MODE == 0) S.k[i] = 0.0f;
    if (MODE == 1) S.k[i] = (float)((double)kold + (double)kerr * (double)kal);
    if (MODE == 2) S.kv[i] = (float)((double)kold + (double)kerr * (double)kal);
}

// Density or divergence error sweep with the scene reduction and the loop
// test.  MODE 0/1: predictDensities (src/DFSPHSolver.cpp:238-309),
//   rho*_i = max(rho_i + dt (sum_j m (v_i - v_j).gradW_ij + v_i.gb_i), rho0),
//   kappa_i = (rho*_i - rho0) alpha_i / (dt^2 rho_i)   (:346)
// MODE 2/3: computeDivergenceError (:408-492),
//   e_i = sum_j m (v_i - v_j).gradW_ij + v_i.gb_i, 0 below 20 nonzero-gradient
//   neighbors, else max(e_i, 0);  kappa_v_i = e_i alpha_i / (dt rho_i) (:531)
// MODE 0/2 is the call before the loop: the loop's error starts at 0
// (:326, :511); 1/3 is the call at the end of an iteration.  Loop test
// (:327, :512; the divergence loop is bounded by maxIterations):
//   continue while (err > eta (eta_V/dt) or it < minIt) and it < maxIt.
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_df_err(const DevScene* __restrict__ T, const Chunk* __restrict__ C) {
    const Chunk c = C[blockIdx.x];
    const DevScene& S = T[c.scene];
    const Scalars* ss = S.ss;
    constexpr bool DENS = (MODE <= 1);
    if (MODE == 1 && ss->done) return;
    if (MODE == 3 && ss->doneV) return;
    __shared__ double sh[kWaves];
    const int i = c.begin + threadIdx.x;
    const KCoefD k = S.kd;
    const double dt = ss->dt;
    double e = 0.0;
    if (i < S.n) {
        const float4 xi = ld4(S.x, i);
        const double xx = xi.x, xy = xi.y, xz = xi.z;
        const float4 gb = S.gb[i];
        // loaded before any store (a wait for a load also waits for the
        // stores issued before it)
        const float drhof = S.drho[i], alf = S.alpha[i];
        const int ngr = DENS ? 20 : (int)S.ngrad[i];
        double dc;
        if (DENS) {
            // the density solve: x.w and vyz hold u, v holds v0 (see store_u)
            const float2 yz = ld2o(S.vyz, (uint32_t)i);
            const double ux = xi.w, uy = yz.x, uz = yz.y;
            double su = 0.0;
            if (MODE == 0 && !ss->d0Valid) {
                // the constant part D0 over v0, once per step (unless the
                // previous divergence sweep left it: see d0Valid)
                const float4 vi = ld4(S.v, i);
                const double wx = vi.x, wy = vi.y, wz = vi.z;
                double s0 = 0.0;
                sweep4(S, i,
                       [&](uint32_t j) { XVU r; r.x = ld4o(S.x, j); r.v = ld4o(S.v, j); r.yz = ld2o(S.vyz, j); return r; },
                       [&](uint32_t, const XVU& d) {
                           const PairG p = pairg_of(xx, xy, xz, d.x, k);
                           s0 += p.g * ((wx - (double)d.v.x) * p.dx + (wy - (double)d.v.y) * p.dy + (wz - (double)d.v.z) * p.dz);
                           su += p.g * ((ux - (double)d.x.w) * p.dx + (uy - (double)d.yz.x) * p.dy + (uz - (double)d.yz.y) * p.dz);
                       });
                const double d0 = S.mD * s0 + (wx * (double)gb.x + wy * (double)gb.y + wz * (double)gb.z);
                ((gptr<double>)S.d0)[i] = d0;
                dc = d0;
                // (d0Valid stays 0 until the next divergence sweep: other
                // blocks of this launch test it)
            } else {
                sweep4(S, i, [&](uint32_t j) { XVP r; r.xv = ld4o(S.x, j); r.yz = ld2o(S.vyz, j); return r; },
                       [&](uint32_t, const XVP& d) {
                           const PairG p = pairg_of(xx, xy, xz, d.xv, k);
                           su += p.g * ((ux - (double)d.xv.w) * p.dx + (uy - (double)d.yz.x) * p.dy + (uz - (double)d.yz.y) * p.dz);
                       });
                dc = ((gptr<const double>)S.d0)[i];
            }
            dc += S.mD * su + (ux * (double)gb.x + uy * (double)gb.y + uz * (double)gb.z);
        } else {
            const float4 vi = ld4(S.v, i);
            const double ux = vi.x, uy = vi.y, uz = vi.z;
            dc = 0.0;
            // the velocity sweep before this one wrote x.w and vyz
            sweep4(S, i, [&](uint32_t j) { XVP r; r.xv = ld4o(S.x, j); r.yz = ld2o(S.vyz, j); return r; },
                   [&](uint32_t, const XVP& d) {
                       const PairG p = pairg_of(xx, xy, xz, d.xv, k);
                       dc += p.g * ((ux - (double)d.xv.w) * p.dx + (uy - (double)d.yz.x) * p.dy + (uz - (double)d.yz.y) * p.dz);
                   });
            dc = S.mD * dc + (ux * (double)gb.x + uy * (double)gb.y + uz * (double)gb.z);
            // the divergence of the current velocities: the next step's D0
            ((gptr<double>)S.d0)[i] = dc;
        }
        const double drho = (double)drhof;
        const double rho = S.rho0D + drho;
        const double al = alf;
        double s;
        if (DENS) {
            const double drs = fmax(drho + dc * dt, 0.0);
            S.drhoS[i] = (float)drs;
            s = drs * al / (dt * dt * rho);
            e = drs / S.rho0D;
        } else {
            const double de = (ngr < 20) ? 0.0 : fmax(dc, 0.0);
            S.divErr[i] = (float)de;
            s = de * al / (dt * rho);
            e = de / S.rho0D;
        }
        S.xk[i] = make_float4(xi.x, xi.y, xi.z, (float)s);
    }
    const double bs = block_sum(e, sh);
    if (threadIdx.x == 0) S.partial[c.begin / kBlock] = bs;
}

// updatePositions (src/DFSPHSolver.cpp:216-236) fused with the cell key and
// histogram of the grid rebuild.
// Also clears the loop counters (scenes iterating, largest count) for the
// divergence solve, which the list build's finalize starts counting.
__global__ __launch_bounds__(kBlock) void k_df_positions(const DevScene* __restrict__ T, const Chunk* __restrict__ C,
                                                         int* active) {
    const Chunk c = C[blockIdx.x];
    const DevScene& S = T[c.scene];
    if (blockIdx.x == 0 && threadIdx.x < 2) active[threadIdx.x] = 0;
    // the lists about to be built start without an overflow
    if (c.begin == 0 && threadIdx.x == 0) G(S.ss)->listOver = 0u;
    const int i = c.begin + threadIdx.x;
    int cl = -1;
    if (i < S.n) {
        const double dt = G(S.ss)->dt;
        float4 x = ld4(S.x, (size_t)i);
        // the density solve's velocity: v0 + u (see store_u)
        const float4 v0 = ld4(S.v, (size_t)i);
        const float2 uyz = ld2o(S.vyz, (uint32_t)i);
        const float4 v = make_float4((float)((double)v0.x + (double)x.w), (float)((double)v0.y + (double)uyz.x),
                                     (float)((double)v0.z + (double)uyz.y), 0.f);
        st4(S.v, (size_t)i, v);
        x.x = (float)((double)x.x + (double)v.x * dt);
        x.y = (float)((double)x.y + (double)v.y * dt);
        x.z = (float)((double)x.z + (double)v.z * dt);
        st4(S.x, (size_t)i, x);
        if (!S.pbOn) {
            cl = cell_of(S, x);
            G(S.cellU)[i] = cl;
        }
    }
    // with the PCISPH boundary method the projection moves the particle
    // first, then counts its cell (a scene-uniform branch)
    if (!S.pbOn) (void)wave_count_add(S.cellCount, cl);
}

// =============================================================== PCISPH

// Predicted densities at x* with the candidate cells of x
// (src/PCISPHSolver.cpp:164-212; every candidate of the 27 cells, self
// included), then updatePressure (:249-295): p_i += delta max(0, rho_i -
// rho0); the scene's max relative error decides the loop (:365), counting
// this iteration as done (its pressure acceleration always follows).
__global__ __launch_bounds__(kBlock) void k_pci_density(const DevScene* __restrict__ T, const Chunk* __restrict__ C) {
    const Chunk c = C[blockIdx.x];
    const DevScene& S = T[c.scene];
    const Scalars* ss = S.ss;
    if (ss->done) return;
    __shared__ double sh[kWaves];
    const int i = c.begin + threadIdx.x;
    const KCoefD k = S.kd;
    const bool useShell = S.pciShell && ss->shellBad == 0u;
    double e = 0.0;
    double dd = 0.0;   // squared predicted displacement, for the next step's shell width
    if (i < S.n) {
        // x* = x + d with d = v* dt held separately: x*_i - x*_j =
        // (x_i - x_j) + (d_i - d_j) keeps FP64 accuracy from FP32 storage
        const float4 xi = S.x[i];
        const float4 di = S.xs[i];
        dd = (double)(di.x * di.x + di.y * di.y + di.z * di.z);
        const double px = xi.x, py = xi.y, pz = xi.z;
        const double ex = di.x, ey = di.y, ez = di.z;
        const int ci = S.cell[i];
        const int nx = S.nx, nxny = S.nx * S.ny;
        double ws = 0.0;
        auto term = [&](const XV& q) {
            const double dx = (px - (double)q.x.x) + (ex - (double)q.v.x);
            const double dy = (py - (double)q.x.y) + (ey - (double)q.v.y);
            const double dz = (pz - (double)q.x.z) + (ez - (double)q.v.z);
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 <= k.hhi2) ws += W_d(sqrt(r2) * k.invh, k);
        };
        auto xv = [&](uint32_t j) { XV r; r.x = ld4o(S.x, j); r.v = ld4o(S.xs, j); return r; };
        const int nsi = useShell ? (int)S.nsh[i] : 0xFFFF;
        if (nsi != 0xFFFF) {
            // every candidate of the 27 cells that can come within h at x*
            // is the particle itself, a list member (|x_i - x_j| <= h) or a
            // shell member (<= h + skin): the same terms in another order
            ws = W_d(0.0, k);
            const uint32_t* L = list_of(S, i);
            const int own = S.nnb[i];
            for (int q0 = 0; q0 < own; q0 += 4) {
                const v4u b = ld_block(L, q0 >> 2);
                const uint32_t jj[4] = {b.x, q0 + 1 < own ? b.y : b.x, q0 + 2 < own ? b.z : b.x,
                                        q0 + 3 < own ? b.w : b.x};
                XV d[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) d[u] = xv(jj[u]);
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (q0 + u < own) term(d[u]);
            }
            // shell member s is in slot maxN - 1 - s: block maxN/4 - 1 - s/4,
            // component 3 - s%4
            const int maxN = S.maxNbr;
            for (int s0 = 0; s0 < nsi; s0 += 4) {
                const v4u b = ld_block(L, (maxN >> 2) - 1 - (s0 >> 2));
                const uint32_t jj[4] = {b.w, s0 + 1 < nsi ? b.z : b.w, s0 + 2 < nsi ? b.y : b.w,
                                        s0 + 3 < nsi ? b.x : b.w};
                XV d[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) d[u] = xv(jj[u]);
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (s0 + u < nsi) term(d[u]);
            }
        } else {
            walk27x4(S.cellStart, ci, nx, nxny, [&](int j) { return xv((uint32_t)j); },
                     [&](int, const XV& q) { term(q); });
        }
        double bw = 0.0;
        walk27x4(S.bCellStart, ci, nx, nxny, [&](int b) { return ld4(S.xb, b); },
                 [&](int, const float4& y) {
                     const double dx = (px - (double)y.x) + ex, dy = (py - (double)y.y) + ey, dz = (pz - (double)y.z) + ez;
                     const double r2 = dx * dx + dy * dy + dz * dz;
                     if (r2 <= k.hhi2) bw += (double)y.w * W_d(sqrt(r2) * k.invh, k);
                 });
        const double rho0 = S.rho0D;
        const double drho = (S.mD * ws + rho0 * bw) - rho0;
        const double rho = rho0 + drho;
        S.drho[i] = (float)drho;
        const double err = fmax(0.0, drho);
        const double pr = (double)S.p[i] + S.pciDeltaD * err;
        S.p[i] = (float)pr;
        S.xk[i] = make_float4(xi.x, xi.y, xi.z, (float)(pr / (rho * rho)));
        e = err / rho0;
    }
    const double bm = block_max(e, sh);
    if (threadIdx.x == 0) S.partial[c.begin / kBlock] = bm;
    if (S.pciShell) {
        // the max is order-free: one atomic per block on non-negative FP32 bits
        const double dm = block_max(dd, sh);
        if (threadIdx.x == 0) atomicMax(&S.ss->maxD2, __float_as_uint((float)dm));
    }
}

// Pressure acceleration with the Spiky gradient at the current positions and
// predicted densities (src/PCISPHSolver.cpp:297-358), then the prediction of
// the next iteration (:122-149).  Runs after every k_pci_density, including
// the one that ended the loop.
__global__ __launch_bounds__(kBlock) void k_pci_pacc(const DevScene* __restrict__ T, const Chunk* __restrict__ C, int iter) {
    const Chunk c = C[blockIdx.x];
    const DevScene& S = T[c.scene];
    const Scalars* ss = S.ss;
    if (ss->it <= iter) return;   // this scene finished before this iteration
    const int i = c.begin + threadIdx.x;
    if (i >= S.n) return;
    const KCoefD k = S.kd;
    const double dt = ss->dt;
    const float4 xi = S.xk[i];
    const double xx = xi.x, xy = xi.y, xz = xi.z;
    const double pri = xi.w;
    const double m = S.mD;
    double ax = 0.0, ay = 0.0, az = 0.0;
    sweep4(S, i, [&](uint32_t j) { return ld4(S.xk, j); },
           [&](uint32_t, const float4& xj) {
               const Pair p = pair_of(xx, xy, xz, xj, k);
               const double g = m * (pri + (double)xj.w) * Gs_d(p, k);
               ax -= g * p.dx;
               ay -= g * p.dy;
               az -= g * p.dz;
           });
    const float4 gs = S.gbs[i];
    ax -= pri * (double)gs.x;
    ay -= pri * (double)gs.y;
    az -= pri * (double)gs.z;
    if (S.pbOn) {
        // m_b p_b / rho_b^2 Spiky gradW of the PCISPH method
        // (src/PCISPHSolver.cpp:330-340)
        walk27(S.bCellStart, S.cell[i], S.nx, S.nx * S.ny, [&](int b) {
            const float4 y = S.xb[b];
            const Pair p = pair_of(xx, xy, xz, y, k);
            const double g = S.rho0D * (double)y.w * (double)S.bpr[b] * Gs_d(p, k);
            ax -= g * p.dx;
            ay -= g * p.dy;
            az -= g * p.dz;
        });
    }
    S.pacc[i] = make_float4((float)ax, (float)ay, (float)az, 0.f);
    const float4 a = S.a[i];
    const float4 v = S.v[i];
    const double wx = (double)v.x + (((double)a.x + ax) + (double)S.gx) * dt;
    const double wy = (double)v.y + (((double)a.y + ay) + (double)S.gy) * dt;
    const double wz = (double)v.z + (((double)a.z + az) + (double)S.gz) * dt;
    S.xs[i] = make_float4((float)(wx * dt), (float)(wy * dt), (float)(wz * dt), 0.f);
    if (S.pbOn) S.pv[i] = make_float4((float)wx, (float)wy, (float)wz, 0.f);
    shell_guard(S, wx * dt, wy * dt, wz * dt);
}

// PCISPH integrate (src/PCISPHSolver.cpp:444-474): a += g + pacc.
__global__ __launch_bounds__(kBlock) void k_pci_integrate(const DevScene* __restrict__ T, const Chunk* __restrict__ C) {
    const Chunk c = C[blockIdx.x];
    const DevScene& S = T[c.scene];
    const int i = c.begin + threadIdx.x;
    if (c.begin == 0 && threadIdx.x == 0 && time_in_last_kernel(S)) advance_time(S);
    if (i >= S.n) return;
    const double dt = S.ss->dt;
    const float4 a = S.a[i], pa = S.pacc[i];
    const float4 v0 = S.v[i];
    const float4 x0 = S.x[i];
    if (S.pbOn) S.x0s[i] = x0;
    const double vx = (double)v0.x + ((double)a.x + ((double)S.gx + (double)pa.x)) * dt;
    const double vy = (double)v0.y + ((double)a.y + ((double)S.gy + (double)pa.y)) * dt;
    const double vz = (double)v0.z + ((double)a.z + ((double)S.gz + (double)pa.z)) * dt;
    S.v[i] = make_float4((float)vx, (float)vy, (float)vz, 0.f);
    S.x[i] = make_float4((float)((double)x0.x + vx * dt), (float)((double)x0.y + vy * dt),
                         (float)((double)x0.z + vz * dt), 0.f);
}

// =========================================================== finalize

// Scene reductions and loop control, one block per scene, after the sweep
// that wrote the chunk partials: the partials are added (or maxed) in chunk
// order, so the result does not depend on scheduling.
//   FIN_DT    CFL time step (src/DFSPHSolver.cpp:827-870): maxVel =
//             max(0.1, max_i sqrt(|v_i + (a_i + g) dt|)); dt' = cfl if
//             cfl < dt, else 0.99 dt after more than 3 density iterations,
//             else 1.001 dt; clamped to [min, max].
//   FIN_DENS0 / FIN_DENS   density loop of src/DFSPHSolver.cpp:320-398
//   FIN_DIV0 / FIN_DIV     divergence loop of src/DFSPHSolver.cpp:505-582
//             (bounded by maxIterations, :512; time += dt at its start)
//   FIN_PCI   PCISPH loop test after updatePressure (src/PCISPHSolver.cpp:365)
// The *0 kinds follow the call before the loop, whose error only feeds the
// log: the loop's error starts at 0 (:326, :511).  active[0] counts the
// scenes still iterating, active[1] is the largest iteration count reached.
enum FinKind { FIN_DT = 0, FIN_DENS0, FIN_DENS, FIN_DIV0, FIN_DIV, FIN_PCI };

template <int KIND>
__global__ __launch_bounds__(kBlock) void k_finalize(const DevScene* __restrict__ T, int nscenes, int* active) {
    const int sc = blockIdx.x;
    if (sc >= nscenes) return;
    const DevScene& S = T[sc];
    const gptr<Scalars> ss = G(S.ss);
    if ((KIND == FIN_DENS || KIND == FIN_PCI) && ss->done) return;
    if (KIND == FIN_DIV && ss->doneV) return;
    __shared__ double sh[kWaves];
    const double red = (KIND == FIN_DT || KIND == FIN_PCI) ? scene_partial_max(S.partial, S.nChunks, sh)
                                                           : scene_partial_sum(S.partial, S.nChunks, sh);
    if (threadIdx.x != 0) return;
    if (KIND == FIN_DT) {
        // the density solve's loop counters start here (counted by FIN_DENS0)
        if (sc == 0) {
            active[0] = 0;
            active[1] = 0;
        }
        const double maxVel = fmax(0.1, red);
        const double ts = ss->dt;
        const double cflTs = S.cfl * 0.4 * 2.0 * S.r / maxVel;
        double nts;
        if (cflTs < ts) nts = cflTs;
        else if (ss->it > 3) nts = ts * 0.99;
        else nts = ts * 1.001;
        nts = fmax(nts, S.dtMin);
        nts = fmin(nts, S.dtMax);
        ss->dt = nts;
        ss->maxVel = maxVel;
        ss->cflTs = cflTs;
        ss->steps += 1u;
    } else if (KIND == FIN_PCI) {
        ss->maxDensErr = red;
        const int it = ++ss->it;
        const bool more = ((red > S.eta) || (it < S.minIt)) && (it < S.maxIt);
        if (!more) {
            ss->done = 1;
            ss->sumIt += (unsigned)it;
            atomicSub(active, 1);
            atomicMax(active + 1, it);
        }
    } else {
        const double avg = (S.n == 0) ? 0.0 : red / (double)S.n;
        const bool dens = (KIND == FIN_DENS0 || KIND == FIN_DENS);
        const bool first = (KIND == FIN_DENS0 || KIND == FIN_DIV0);
        int it;
        double err;
        if (first) {
            it = 0;
            err = 0.0;
            if (dens) {
                ss->avgErr0 = avg;
                ss->it = 0;
            } else {
                ss->avgErrV0 = avg;
                ss->itV = 0;
                ss->time += ss->dt;
            }
        } else {
            err = avg;
            if (dens) {
                it = ++ss->it;
                ss->avgErr = avg;
            } else {
                it = ++ss->itV;
                ss->avgErrV = avg;
            }
        }
        const bool more = dens ? (((err > S.eta) || (it < S.minIt)) && (it < S.maxIt))
                               : (((err > S.etaV / ss->dt) || (it < S.minItV)) && (it < S.maxIt));
        if (dens) ss->done = more ? 0 : 1;
        else ss->doneV = more ? 0 : 1;
        if (first && more) atomicAdd(active, 1);
        if (!first && !more) atomicSub(active, 1);
        if (!more) atomicMax(active + 1, it);
        if (!more) {
            if (dens) ss->sumIt += (unsigned)it;
            else ss->sumItV += (unsigned)it;
        }
    }
}

// CubeBoundaryModel::correctPositionAndVelocity (src/CubeBoundaryModel.cpp:
// 16-94): per axis in x, y, z order, a position below the minimum (above
// the maximum) is set to it, the axis velocity is multiplied by -normalFct if
// it points out of the box, and the other two components by tangentialFct.
// The faces are the nearest FP32 values inside the box, so the set of
// clamped particles is the FP64 comparison's, and a particle left on a face
// is not clamped again on later steps (FP64 positions sit exactly on it).
__global__ __launch_bounds__(kBlock) void k_cube(const DevScene* __restrict__ T, const Chunk* __restrict__ C) {
    const Chunk c = C[blockIdx.x];
    const DevScene& S = T[c.scene];
    const int i = c.begin + threadIdx.x;
    if (!S.cubeOn) return;
    // the clamp changes velocities after the divergence sweeps formed d0
    if (c.begin == 0 && threadIdx.x == 0 && S.method == DFSPH) S.ss->d0Valid = 0u;
    if (i >= S.n) return;
    const float4 xf = S.x[i], vf = S.v[i];
    float p[3] = {xf.x, xf.y, xf.z};
    double u[3] = {vf.x, vf.y, vf.z};
    const double nf = S.cubeNf, tf = S.cubeTf;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int b = (a + 1) % 3, e = (a + 2) % 3;
        bool hit = false;
        if (p[a] < S.cubeLo[a]) {
            p[a] = S.cubeLo[a];
            if (u[a] < 0.0) u[a] *= -nf;
            hit = true;
        } else if (p[a] > S.cubeHi[a]) {
            p[a] = S.cubeHi[a];
            if (u[a] > 0.0) u[a] *= -nf;
            hit = true;
        }
        if (hit) {
            u[b] *= tf;
            u[e] *= tf;
        }
    }
    S.x[i] = make_float4(p[0], p[1], p[2], xf.w);
    S.v[i] = make_float4((float)u[0], (float)u[1], (float)u[2], vf.w);
}

// Emitters (Simulation::emitParticles -> Emitter::emitParticles,
// src/Emitter.cpp:85-162), one block per scene, emitters in order: an
// emitter whose next emission time has come (time >= next) and whose budget
// holds a whole group appends the group at slots n, n+1, ... (ids equal to
// the slots, continuing the active count): velocity (0, 0, vel) and
// position group + t (velocity + g t), t = time - next, both rotated by the
// quaternion as glm evaluates it (u = (q.x, q.y, q.z):
// v + ((u x v) w + u x (u x v)) 2), the position then offset by the
// emitter origin.  Warm-start values start at zero.  With `cells`, the new
// particles' cells are counted for the grid sort that follows (DFSPH emits
// between its position update and the grid rebuild).  Tc is the table whose
// buffers are current at the call (before or after the step's sort), To
// the other; the new count goes to both and to the scalars.
__device__ inline void quat_rotate(const double q[4], double& x, double& y, double& z) {
    const double ux = q[1], uy = q[2], uz = q[3];
    const double cx = uy * z - y * uz, cy = uz * x - z * ux, cz = ux * y - x * uy;      // u x v
    const double dx = uy * cz - cy * uz, dy = uz * cx - cz * ux, dz = ux * cy - cx * uy;  // u x (u x v)
    x = x + ((cx * q[0]) + dx) * 2.0;
    y = y + ((cy * q[0]) + dy) * 2.0;
    z = z + ((cz * q[0]) + dz) * 2.0;
}

__global__ __launch_bounds__(kBlock) void k_emit(DevScene* __restrict__ Tc, DevScene* __restrict__ To, int nscenes,
                                                 int cells) {
    const int sc = blockIdx.x;
    if (sc >= nscenes) return;
    const DevScene& S = Tc[sc];
    if (S.nEmit == 0) return;
    Scalars* ss = S.ss;
    const double time = ss->time;
    const int n0 = S.n;
    int n = n0;
    for (int e = 0; e < S.nEmit; ++e) {
        EmitterDev* E = S.emit + e;
        const int gs = E->groupSize;
        const bool fire = (time >= E->next) && (E->remaining >= (unsigned)gs) && (n + gs <= S.cap);
        if (fire) {
            const double t = time - E->next;
            double q[4] = {E->q[0], E->q[1], E->q[2], E->q[3]};
            for (int g = threadIdx.x; g < gs; g += blockDim.x) {
                const double* lp = S.emitGroup + 3 * (size_t)(E->group + g);
                double vx = 0.0, vy = 0.0, vz = E->vel;
                double px = lp[0] + (vx + (double)S.gx * t) * t;
                double py = lp[1] + (vy + (double)S.gy * t) * t;
                double pz = lp[2] + (vz + (double)S.gz * t) * t;
                quat_rotate(q, px, py, pz);
                quat_rotate(q, vx, vy, vz);
                px += E->pos[0];
                py += E->pos[1];
                pz += E->pos[2];
                const int slot = n + g;
                const float4 xp = make_float4((float)px, (float)py, (float)pz, 0.f);
                S.x[slot] = xp;
                S.v[slot] = make_float4((float)vx, (float)vy, (float)vz, 0.f);
                S.id[slot] = (uint32_t)slot;
                if (S.k) {
                    S.k[slot] = 0.f;
                    S.kv[slot] = 0.f;
                }
                if (cells) {
                    const int cl = cell_of(S, xp);
                    S.cellU[slot] = cl;
                    atomicAdd(&S.cellCount[cl], 1);
                }
            }
            n += gs;
        }
        __syncthreads();   // every thread has read this emitter's state
        if (fire && threadIdx.x == 0) {
            E->remaining -= (unsigned)gs;
            E->next += E->interval;
        }
    }
    if (threadIdx.x == 0 && n != n0) {
        Tc[sc].n = n;
        To[sc].n = n;
        ss->nActive = (unsigned)n;
    }
}

// ===================================================== PCISPH boundary method

// The boundary particles' sums over the (static) boundary, m_b W and m_b
// Poly6 with the boundary candidates of the particle's cell
// (src/SPHSolver.cpp:115-120, src/PCISPHSolver.cpp:239-244).  One scene.
__global__ __launch_bounds__(kBlock) void k_pb_bself(const DevScene* __restrict__ T, int scene) {
    const DevScene& S = T[scene];
    const int b = blockIdx.x * kBlock + threadIdx.x;
    if (b >= S.nb) return;
    const KCoefD k = S.kd;
    const float4 yb = S.xb[b];
    const double px = yb.x, py = yb.y, pz = yb.z;
    double wc = 0.0, wp = 0.0;
    walk27(S.bCellStart, cell_of(S, yb), S.nx, S.nx * S.ny, [&](int c) {
        const float4 y = S.xb[c];
        const Pair p = pair_of(px, py, pz, y, k);
        const double mb = S.rho0D * (double)y.w;
        wc += mb * W_d(p.q, k);
        wp += mb * P6_d(p.r2, k);
    });
    S.bself[2 * b] = wc;
    S.bself[2 * b + 1] = wp;
}

// Boundary densities and pressures (one scene).  PRED = 0, WCSPH
// (src/SPHSolver.cpp:93-121, src/WCSPHSolver.cpp:97-114): rho_b = sum over
// the fluid candidates of the particle's cell of m W(r_b - x_j) plus the
// boundary sum, p_b = max(0, kappa rho0/gamma ((rho_b/rho0)^gamma - 1)).
// PRED = 1, PCISPH loop (src/PCISPHSolver.cpp:214-247, :276-293): Poly6
// with the predicted fluid positions, p_b += delta max(0, rho_b - rho0),
// from zero at the first iteration.  Both store p_b / rho_b^2 for the
// fluid's pressure sums.  In the loop it runs after the finalize of
// iteration `iter`, for scenes that ran that iteration.
template <int PRED>
__global__ __launch_bounds__(kBlock) void k_pb_bdensity(const DevScene* __restrict__ T, int scene, int iter) {
    const DevScene& S = T[scene];
    if (PRED && S.ss->it <= iter) return;
    const int first = iter == 0;
    const int b = blockIdx.x * kBlock + threadIdx.x;
    if (b >= S.nb) return;
    const KCoefD k = S.kd;
    const float4 yb = S.xb[b];
    const double px = yb.x, py = yb.y, pz = yb.z;
    const double m = S.mD, rho0 = S.rho0D;
    double s = 0.0;
    walk27(S.cellStart, cell_of(S, yb), S.nx, S.nx * S.ny, [&](int j) {
        const float4 xj = S.x[j];
        if (PRED) {
            const float4 dj = S.xs[j];
            const double dx = (px - (double)xj.x) - (double)dj.x;
            const double dy = (py - (double)xj.y) - (double)dj.y;
            const double dz = (pz - (double)xj.z) - (double)dj.z;
            s += m * P6_d(dx * dx + dy * dy + dz * dz, k);
        } else {
            const Pair p = pair_of(px, py, pz, xj, k);
            s += m * W_d(p.q, k);
        }
    });
    const double rho = s + S.bself[2 * b + PRED];
    double pr;
    if (PRED) {
        pr = (first ? 0.0 : (double)S.bpres[b]) + S.pciDeltaD * fmax(0.0, rho - rho0);
    } else {
        const double x = (rho - rho0) / rho0;
        pr = fmax(0.0, (double)S.stiffness * rho0 / (double)S.gammaEos * expm1((double)S.gammaEos * log1p(x)));
    }
    S.bdens[b] = (float)rho;
    S.bpres[b] = (float)pr;
    S.bpr[b] = (float)(pr / (rho * rho));
}

// Position (and velocity) projection after the step's position update.
// MODE 0, DFSPH (src/DFSPHSolver.cpp:65-75): q = x; the particle's cell is
// then counted for the grid rebuild.  MODE 1, WCSPH (src/WCSPHSolver.cpp:
// 62-73): q = x, then the velocity.  MODE 2, PCISPH (src/PCISPHSolver.cpp:
// 65-76): q = the loop's last corrected predicted position, and the velocity
// correction starts from its predicted velocity.
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_pb_correct(const DevScene* __restrict__ T, const Chunk* __restrict__ C) {
    const Chunk c = C[blockIdx.x];
    const DevScene& S = T[c.scene];
    if (!S.pbOn) return;
    const int i = c.begin + threadIdx.x;
    if (i >= S.n) return;
    float4 xf = S.x[i];
    const double px = xf.x, py = xf.y, pz = xf.z;
    double qx = px, qy = py, qz = pz;
    if (MODE == 2) {
        const float4 x0 = S.x0s[i], d = S.prS[i];
        qx = (double)x0.x + (double)d.x;
        qy = (double)x0.y + (double)d.y;
        qz = (double)x0.z + (double)d.z;
    }
    const PbHit h = pb_project(S, px, py, pz, xf, qx, qy, qz);
    S.pbn[i] = make_float4((float)h.nx, (float)h.ny, (float)h.nz, 0.f);
    if (h.hit) {
        xf.x = (float)(px + h.s * h.nx);
        xf.y = (float)(py + h.s * h.ny);
        xf.z = (float)(pz + h.s * h.nz);
        S.x[i] = xf;
    }
    if (MODE == 0) {
        const int cl = cell_of(S, xf);
        S.cellU[i] = cl;
        atomicAdd(&S.cellCount[cl], 1);
    } else {
        double ux, uy, uz;
        if (MODE == 1) {
            const float4 v = S.v[i];
            ux = v.x; uy = v.y; uz = v.z;
        } else {
            const float4 v = S.pvS[i];
            ux = v.x; uy = v.y; uz = v.z;
        }
        if (pb_velocity(S, h.nx, h.ny, h.nz, ux, uy, uz)) S.v[i] = make_float4((float)ux, (float)uy, (float)uz, 0.f);
    }
}

// correctPredPositions (src/PCISPHBoundaryModel.cpp:385-439), each PCISPH
// loop iteration after the prediction: collision test with the current
// position, weights and shift on the predicted one (x + d), d corrected.
// Keeps the corrected d and the predicted velocity for the projection after
// the integration.
__global__ __launch_bounds__(kBlock) void k_pb_predcorrect(const DevScene* __restrict__ T, const Chunk* __restrict__ C) {
    const Chunk c = C[blockIdx.x];
    const DevScene& S = T[c.scene];
    if (!S.pbOn || S.ss->done) return;
    const int i = c.begin + threadIdx.x;
    if (i >= S.n) return;
    const float4 xf = S.x[i];
    float4 d = S.xs[i];
    const double px = xf.x, py = xf.y, pz = xf.z;
    const PbHit h = pb_project(S, px, py, pz, xf, px + (double)d.x, py + (double)d.y, pz + (double)d.z);
    S.pbn[i] = make_float4((float)h.nx, (float)h.ny, (float)h.nz, 0.f);
    if (h.hit) {
        d.x = (float)((double)d.x + h.s * h.nx);
        d.y = (float)((double)d.y + h.s * h.ny);
        d.z = (float)((double)d.z + h.s * h.nz);
        S.xs[i] = d;
    }
    S.prS[i] = d;
    S.pvS[i] = S.pv[i];
}

// WCSPH: time += dt after the integration (src/WCSPHSolver.cpp:77).
__global__ void k_time_advance(const DevScene* __restrict__ T, int nscenes) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nscenes || time_in_last_kernel(T[s])) return;
    advance_time(T[s]);
}

// The neighbor-overflow maximum of every scene of a batch starts each step
// call at the overflow of the lists the call starts with -- DFSPH's first
// sweeps use the lists of the previous call's last rebuild (or of init /
// set_state); WCSPH and PCISPH build theirs at the start of each step -- so
// it reports the lists of that call's steps only.
__global__ void k_clear_overflow(const DevScene* __restrict__ T, int nscenes) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nscenes) T[s].ss->overflow = (T[s].method == DFSPH) ? T[s].ss->listOver : 0u;
}

// ============================================================ utilities

__global__ void k_fill_int(int* p, int v, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Sorted -> particle-id order.
__global__ void k_unsort_f4(const float4* __restrict__ src, const uint32_t* __restrict__ id, float* __restrict__ dst, int n, int comps) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const float4 v = src[s];
    float* d = dst + (size_t)id[s] * comps;
    d[0] = v.x;
    if (comps > 1) d[1] = v.y;
    if (comps > 2) d[2] = v.z;
    if (comps > 3) d[3] = v.w;
}
__global__ void k_unsort_f1(const float* __restrict__ src, const uint32_t* __restrict__ id, float* __restrict__ dst, int n,
                            double offset) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    dst[id[s]] = (float)(offset + (double)src[s]);
}
__global__ void k_unsort_i1(const int* __restrict__ src, const uint32_t* __restrict__ id, int* __restrict__ dst, int n) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    dst[id[s]] = src[s];
}
// Reference keys floor(x/h) of each particle, in id order.
__global__ void k_keys3(const DevScene* __restrict__ T, int scene, int* __restrict__ dst) {
    const DevScene& S = T[scene];
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S.n) return;
    const float4 p = S.x[s];
    int* d = dst + (size_t)S.id[s] * 3;
    d[0] = (int)floor(__ddiv_rn((double)p.x, S.hD));
    d[1] = (int)floor(__ddiv_rn((double)p.y, S.hD));
    d[2] = (int)floor(__ddiv_rn((double)p.z, S.hD));
}
// Id-ordered upload -> working arrays (identity order until the first sort).
__global__ void k_load_state(const DevScene* __restrict__ T, int scene, const float* __restrict__ xyz, const float* __restrict__ vel,
                             const float* __restrict__ kk, const float* __restrict__ kkv) {
    const DevScene& S = T[scene];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S.n) return;
    S.x[i] = make_float4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], 0.f);
    S.v[i] = vel ? make_float4(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2], 0.f) : make_float4(0.f, 0.f, 0.f, 0.f);
    S.id[i] = (uint32_t)i;
    if (S.a) S.a[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (S.k) S.k[i] = kk ? kk[i] : 0.0f;
    if (S.kv) S.kv[i] = kkv ? kkv[i] : 0.0f;
}

// ============================================================ launchers

#define LAUNCH(kern, grid, stream, ...) hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), 0, stream, __VA_ARGS__)

void launch_keys(hipStream_t st, const DevScene* T, const Chunk* C, int n, int* begin_active) {
    if (!n) return;
    if (begin_active) LAUNCH(k_keys<true>, n, st, T, C, begin_active);
    else LAUNCH(k_keys<false>, n, st, T, C, nullptr);
}
void launch_scan(hipStream_t st, const DevScene* T, const Chunk* CC, int ncc, int nscenes, int maxChunks) {
    if (!ncc) return;
    LAUNCH(k_scan_sum, ncc, st, T, CC);
    if (maxChunks <= kScanFusedChunks) {
        LAUNCH(k_scan_final<true>, ncc, st, T, CC);
        return;
    }
    LAUNCH(k_scan_offsets, nscenes, st, T, nscenes);
    LAUNCH(k_scan_final<false>, ncc, st, T, CC);
}
void launch_scatter(hipStream_t st, const DevScene* T, const Chunk* C, int n) { if (n) LAUNCH(k_scatter, n, st, T, C); }
void launch_rank_reorder(hipStream_t st, const DevScene* T, const Chunk* C, int n) { if (n) LAUNCH(k_rank_reorder, n, st, T, C); }
void launch_bvolume(hipStream_t st, const DevScene* T, int scene, int nb, float* tmp) {
    if (!nb) return;
    const int g = (nb + kBlock - 1) / kBlock;
    LAUNCH(k_bvolume, g, st, T, scene, tmp);
    LAUNCH(k_bvolume_pack, g, st, T, scene, (const float*)tmp);
}
void launch_build(hipStream_t st, int method, const DevScene* T, const Chunk* C, int n) {
    if (!n) return;
    if (method == DFSPH) LAUNCH(k_build<DFSPH>, n, st, T, C);
    else if (method == WCSPH) LAUNCH(k_build<WCSPH>, n, st, T, C);
    else LAUNCH(k_build<PCISPH>, n, st, T, C);
}
void launch_normals(hipStream_t st, const DevScene* T, const Chunk* C, int n) { if (n) LAUNCH(k_normals, n, st, T, C); }
void launch_forces(hipStream_t st, int method, const DevScene* T, const Chunk* C, int n) {
    if (!n) return;
    if (method == DFSPH) LAUNCH(k_forces<DFSPH>, n, st, T, C);
    else if (method == WCSPH) LAUNCH(k_forces<WCSPH>, n, st, T, C);
    else LAUNCH(k_forces<PCISPH>, n, st, T, C);
}
void launch_dt(hipStream_t st, const DevScene* T, const Chunk* C, int n) { if (n) LAUNCH(k_dt, n, st, T, C); }
void launch_df_predict(hipStream_t st, const DevScene* T, const Chunk* C, int n) { if (n) LAUNCH(k_df_predict, n, st, T, C); }
void launch_df_vel(hipStream_t st, int mode, const DevScene* T, const Chunk* C, int n) {
    if (!n) return;
    if (mode == 0) LAUNCH(k_df_vel<0>, n, st, T, C);
    else if (mode == 1) LAUNCH(k_df_vel<1>, n, st, T, C);
    else LAUNCH(k_df_vel<2>, n, st, T, C);
}
void launch_df_err(hipStream_t st, int mode, const DevScene* T, const Chunk* C, int n) {
    if (!n) return;
    if (mode == 0) LAUNCH(k_df_err<0>, n, st, T, C);
    else if (mode == 1) LAUNCH(k_df_err<1>, n, st, T, C);
    else if (mode == 2) LAUNCH(k_df_err<2>, n, st, T, C);
    else LAUNCH(k_df_err<3>, n, st, T, C);
}
void launch_finalize(hipStream_t st, int kind, const DevScene* T, int nscenes, int* active) {
    if (!nscenes) return;
    switch (kind) {
        case FIN_DT: LAUNCH(k_finalize<FIN_DT>, nscenes, st, T, nscenes, active); break;
        case FIN_DENS0: LAUNCH(k_finalize<FIN_DENS0>, nscenes, st, T, nscenes, active); break;
        case FIN_DENS: LAUNCH(k_finalize<FIN_DENS>, nscenes, st, T, nscenes, active); break;
        case FIN_DIV0: LAUNCH(k_finalize<FIN_DIV0>, nscenes, st, T, nscenes, active); break;
        case FIN_DIV: LAUNCH(k_finalize<FIN_DIV>, nscenes, st, T, nscenes, active); break;
        default: LAUNCH(k_finalize<FIN_PCI>, nscenes, st, T, nscenes, active); break;
    }
}
void launch_df_positions(hipStream_t st, const DevScene* T, const Chunk* C, int n, int* active) {
    if (n) LAUNCH(k_df_positions, n, st, T, C, active);
}
void launch_pci_density(hipStream_t st, const DevScene* T, const Chunk* C, int n) { if (n) LAUNCH(k_pci_density, n, st, T, C); }
void launch_pci_pacc(hipStream_t st, const DevScene* T, const Chunk* C, int n, int iter) { if (n) LAUNCH(k_pci_pacc, n, st, T, C, iter); }
void launch_pci_integrate(hipStream_t st, const DevScene* T, const Chunk* C, int n) { if (n) LAUNCH(k_pci_integrate, n, st, T, C); }
void launch_cube(hipStream_t st, const DevScene* T, const Chunk* C, int n) { if (n) LAUNCH(k_cube, n, st, T, C); }
void launch_pb_bself(hipStream_t st, const DevScene* T, int scene, int nb) {
    if (nb) hipLaunchKernelGGL(k_pb_bself, dim3((nb + kBlock - 1) / kBlock), dim3(kBlock), 0, st, T, scene);
}
void launch_pb_bdensity(hipStream_t st, int pred, const DevScene* T, int scene, int nb, int iter) {
    if (!nb) return;
    const dim3 g((nb + kBlock - 1) / kBlock), b(kBlock);
    if (pred) hipLaunchKernelGGL(k_pb_bdensity<1>, g, b, 0, st, T, scene, iter);
    else hipLaunchKernelGGL(k_pb_bdensity<0>, g, b, 0, st, T, scene, iter);
}
// dst[s] = src[id[s]] (an id-ordered float3 array into the working order).
__global__ void k_gather_f3(float4* __restrict__ dst, const float* __restrict__ src, const uint32_t* __restrict__ id,
                            int n) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const float* q = src + 3 * (size_t)id[s];
    dst[s] = make_float4(q[0], q[1], q[2], 0.f);
}
void launch_gather_f3(hipStream_t st, float4* dst, const float* src, const uint32_t* id, int n) {
    if (n) hipLaunchKernelGGL(k_gather_f3, dim3((n + 255) / 256), dim3(256), 0, st, dst, src, id, n);
}
void launch_pb_correct(hipStream_t st, int mode, const DevScene* T, const Chunk* C, int n) {
    if (!n) return;
    if (mode == 0) LAUNCH(k_pb_correct<0>, n, st, T, C);
    else if (mode == 1) LAUNCH(k_pb_correct<1>, n, st, T, C);
    else LAUNCH(k_pb_correct<2>, n, st, T, C);
}
void launch_pb_predcorrect(hipStream_t st, const DevScene* T, const Chunk* C, int n) {
    if (n) LAUNCH(k_pb_predcorrect, n, st, T, C);
}
void launch_emit(hipStream_t st, DevScene* Tc, DevScene* To, int nscenes, int cells) {
    if (nscenes) hipLaunchKernelGGL(k_emit, dim3(nscenes), dim3(kBlock), 0, st, Tc, To, nscenes, cells);
}
void launch_time_advance(hipStream_t st, const DevScene* T, int nscenes) {
    hipLaunchKernelGGL(k_time_advance, dim3((nscenes + 255) / 256), dim3(256), 0, st, T, nscenes);
}
void launch_clear_overflow(hipStream_t st, const DevScene* T, int nscenes) {
    hipLaunchKernelGGL(k_clear_overflow, dim3((nscenes + 255) / 256), dim3(256), 0, st, T, nscenes);
}
void launch_fill_int(hipStream_t st, int* p, int v, long n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_int, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, v, n);
}
void launch_unsort_f4(hipStream_t st, const float4* src, const uint32_t* id, float* dst, int n, int comps) {
    if (n) hipLaunchKernelGGL(k_unsort_f4, dim3((n + 255) / 256), dim3(256), 0, st, src, id, dst, n, comps);
}
void launch_unsort_f1(hipStream_t st, const float* src, const uint32_t* id, float* dst, int n, double offset) {
    if (n) hipLaunchKernelGGL(k_unsort_f1, dim3((n + 255) / 256), dim3(256), 0, st, src, id, dst, n, offset);
}
void launch_unsort_i1(hipStream_t st, const int* src, const uint32_t* id, int* dst, int n) {
    if (n) hipLaunchKernelGGL(k_unsort_i1, dim3((n + 255) / 256), dim3(256), 0, st, src, id, dst, n);
}
void launch_keys3(hipStream_t st, const DevScene* T, int scene, int n, int* dst) {
    if (n) hipLaunchKernelGGL(k_keys3, dim3((n + 255) / 256), dim3(256), 0, st, T, scene, dst);
}
void launch_load_state(hipStream_t st, const DevScene* T, int scene, int n, const float* xyz, const float* vel, const float* k, const float* kv) {
    if (n) hipLaunchKernelGGL(k_load_state, dim3((n + 255) / 256), dim3(256), 0, st, T, scene, xyz, vel, k, kv);
}

}  // namespace sphg
