// sphg_kernels.hip -- gfx950 kernels of the SPH solver step.
//
// Every kernel takes a scene table and a chunk table; block b works on
// chunk b (256 particles or 4096 cells of one scene).  Sums of a scene are
// formed in a fixed order: a butterfly inside each wavefront, wave results
// combined in index order, one partial per chunk, and the last chunk to
// finish (an atomic arrival counter) adds the scene's partials in chunk
// order.  Results therefore do not depend on scheduling.
//
// Reference algorithm: josemanuel28v/SPH src/SPHSolver.cpp,
// src/DFSPHSolver.cpp, src/WCSPHSolver.cpp, src/PCISPHSolver.cpp, the kernel
// headers and the force headers; each kernel names the lines it restates.

#include "sphg_internal.h"
#include "sphg_kernels.h"

namespace sphg {

// ------------------------------------------------------------ utilities

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
// Result valid in every thread.  sh: 4 doubles of shared memory.
__device__ inline double block_sum(double v, double* sh) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double r = ((sh[0] + sh[1]) + sh[2]) + sh[3];
    __syncthreads();
    return r;
}
__device__ inline double block_max(double v, double* sh) {
    v = wave_max(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double r = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
    __syncthreads();
    return r;
}

// Arrival counter: returns true in every thread of the last block of the
// scene to arrive.  The caller has written its partial from thread 0.
__device__ inline bool arrive_last(unsigned* ctr, unsigned total) {
    __shared__ int s_last;
    if (threadIdx.x == 0) {
        __threadfence();
        const unsigned prev = atomicAdd(ctr, 1u);
        s_last = (prev == total - 1u);
    }
    __syncthreads();
    const bool last = s_last != 0;
    if (last) __threadfence();
    return last;
}

// Sum (or max) of one scene's chunk partials in chunk order.
__device__ inline double scene_partial_sum(const double* partial, int n, double* sh) {
    double acc = 0.0;
    for (int t = threadIdx.x; t < n; t += kBlock) acc += partial[t];
    return block_sum(acc, sh);
}
__device__ inline double scene_partial_max(const double* partial, int n, double* sh) {
    double acc = -1e300;
    for (int t = threadIdx.x; t < n; t += kBlock) acc = fmax(acc, partial[t]);
    return block_max(acc, sh);
}

// ------------------------------------------------------- SPH kernels, FP32

// Cubic spline W(r) (src/CubicSpline.h:29-48).
__device__ inline float W_cubic(float r2, const KCoef& k) {
    const float rl = sqrtf(r2);
    const float q = rl * k.invh;
    const float om = 1.0f - q;
    const float w = (q <= 0.5f) ? k.ck * ((6.0f * q - 6.0f) * q * q + 1.0f) : 2.0f * k.ck * om * om * om;
    return (q <= 1.0f) ? w : 0.0f;
}
// Cubic spline gradient as a factor of r: gradW(r) = g * r, zero for
// |r| <= 1e-5 or q > 1 (src/CubicSpline.h:55-73).  For q <= 1/2 the
// reference's l q (3q-2) / (|r| h) equals l (3q-2) / h^2.
__device__ inline float G_cubic(float r2, const KCoef& k) {
    const float inv = rsqrtf(r2);
    const float rl = r2 * inv;
    const float q = rl * k.invh;
    const float om = 1.0f - q;
    const float g = (q <= 0.5f) ? k.clh2 * (3.0f * q - 2.0f) : -k.clh * om * om * inv;
    return (rl > 1e-5f && q <= 1.0f) ? g : 0.0f;
}
// Spiky gradient factor (src/Spiky.h:44-55): -45/(pi h^6) (h-r)^2 / r.
__device__ inline float G_spiky(float r2, const KCoef& k) {
    const float inv = rsqrtf(r2);
    const float rl = r2 * inv;
    const float d = k.h - rl;
    const float g = -k.spg * d * d * inv;
    return (r2 > 0.0f && rl <= k.h) ? g : 0.0f;
}
// Cohesion C(r) (src/Cohesion.h:29-40) for 0 < r <= h; `outer` is the
// branch test 2r > h.  As written there the h^6/64 offset is not scaled by
// 32/(pi h^9), so C jumps at r = h/2 and the branch is decided exactly by
// the caller near that point.
__device__ inline float C_cohesion(float rl, bool outer, const KCoef& k) {
    const float d = k.h - rl;
    const float t = d * d * d * rl * rl * rl;
    return outer ? k.co1 * t : k.co1 * 2.0f * t - k.co2;
}
// Adhesion A(r) (src/Adhesion.h:24-35): (-4r^2/h + 6r - 2h)^(1/4) on
// (h/2, h], evaluated as ((2/h)(2r-h)(h-r))^(1/4), the same polynomial in a
// form that cannot round below zero.
__device__ inline float A_adhesion(float rl, const KCoef& k) {
    if (!(2.0f * rl > k.h && rl <= k.h)) return 0.0f;
    const float v = (2.0f * k.invh) * (2.0f * rl - k.h) * (k.h - rl);
    return sqrtf(sqrtf(fmaxf(v, 0.0f)));
}

// Exact FP64 distance test, operation for operation the reference's
// length(ri - rj) / h <= 1 on FP32 positions promoted to FP64
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

__device__ inline const uint32_t* list_of(const DevScene& S, int i) {
    return S.nbr + ((size_t)(i >> 6) * (size_t)S.maxNbr) * kGroup + (i & 63);
}
__device__ inline uint32_t* list_of_w(const DevScene& S, int i) {
    return S.nbr + ((size_t)(i >> 6) * (size_t)S.maxNbr) * kGroup + (i & 63);
}

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

// ================================================================ grid

// Cell key + histogram (src/HashTable.h:79-86 insertion, as a count).
__global__ __launch_bounds__(kBlock) void k_keys(const DevScene* __restrict__ T, const Chunk* __restrict__ C) {
    const Chunk c = C[blockIdx.x];
    const DevScene& S = T[c.scene];
    const int i = c.begin + threadIdx.x;
    if (i >= S.n) return;
    const int cl = cell_of(S, S.x[i]);
    S.cellU[i] = cl;
    atomicAdd(&S.cellCount[cl], 1);
}

// Per 4096-cell chunk: total count.
__global__ __launch_bounds__(kBlock) void k_scan_sum(const DevScene* __restrict__ T, const Chunk* __restrict__ C) {
    const Chunk c = C[blockIdx.x];
    const DevScene& S = T[c.scene];
    __shared__ double sh[4];
    int acc = 0;
    const int c0 = c.begin + threadIdx.x * kCellPerThread;
#pragma unroll
    for (int u = 0; u < kCellPerThread; ++u) {
        const int cc = c0 + u;
        if (cc < S.ncells) acc += S.cellCount[cc];
    }
    const double tot = block_sum((double)acc, sh);
    if (threadIdx.x == 0) S.cellChunkSum[c.begin / kCellChunk] = (int)tot;
}

// One block per scene: exclusive scan of the chunk totals, in place.
__global__ __launch_bounds__(kBlock) void k_scan_offsets(const DevScene* __restrict__ T, int nscenes) {
    const int s = blockIdx.x;
    if (s >= nscenes) return;
    const DevScene& S = T[s];
    if (threadIdx.x == 0) {
        int run = 0;
        for (int t = 0; t < S.nCellChunks; ++t) {
            const int v = S.cellChunkSum[t];
            S.cellChunkSum[t] = run;
            run += v;
        }
    }
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

// Final scan: cellStart[c] = exclusive prefix of counts; counts reset.
__global__ __launch_bounds__(kBlock) void k_scan_final(const DevScene* __restrict__ T, const Chunk* __restrict__ C) {
    const Chunk c = C[blockIdx.x];
    const DevScene& S = T[c.scene];
    __shared__ int wsum[4];
    const int c0 = c.begin + threadIdx.x * kCellPerThread;
    int vals[kCellPerThread];
    int local = 0;
#pragma unroll
    for (int u = 0; u < kCellPerThread; ++u) {
        const int cc = c0 + u;
        vals[u] = (cc < S.ncells) ? S.cellCount[cc] : 0;
        local += vals[u];
    }
    const int incl = wave_incl_scan(local);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) wsum[w] = incl;
    __syncthreads();
    int woff = 0;
    for (int q = 0; q < w; ++q) woff += wsum[q];
    int run = S.cellChunkSum[c.begin / kCellChunk] + woff + incl - local;
#pragma unroll
    for (int u = 0; u < kCellPerThread; ++u) {
        const int cc = c0 + u;
        if (cc < S.ncells) {
            S.cellStart[cc] = run;
            S.cellCount[cc] = 0;
        }
        run += vals[u];
    }
    // the chunk table covers ncells + 1 entries: the last one is the end
    if (c0 <= S.ncells && S.ncells < c0 + kCellPerThread) S.cellStart[S.ncells] = S.n;
}

// Scatter into cell slots (order inside a cell fixed afterwards).
__global__ __launch_bounds__(kBlock) void k_scatter(const DevScene* __restrict__ T, const Chunk* __restrict__ C) {
    const Chunk c = C[blockIdx.x];
    const DevScene& S = T[c.scene];
    const int i = c.begin + threadIdx.x;
    if (i >= S.n) return;
    const int cl = S.cellU[i];
    const int slot = S.cellStart[cl] + atomicAdd(&S.cellCount[cl], 1);
    S.perm[slot] = (uint32_t)i;
}

// Order each cell by particle id: the reference inserts in ascending id
// (src/SPHSolver.cpp:188-191), so a bucket lists its particles by id.
__global__ __launch_bounds__(kBlock) void k_cell_sort(const DevScene* __restrict__ T, const Chunk* __restrict__ C) {
    const Chunk c = C[blockIdx.x];
    const DevScene& S = T[c.scene];
    const int c0 = c.begin + threadIdx.x * kCellPerThread;
    for (int u = 0; u < kCellPerThread; ++u) {
        const int cc = c0 + u;
        if (cc >= S.ncells) break;
        const int s = S.cellStart[cc], e = S.cellStart[cc + 1];
        S.cellCount[cc] = 0;
        for (int a = s + 1; a < e; ++a) {
            const uint32_t pa = S.perm[a];
            const uint32_t ida = S.id[pa];
            int b = a - 1;
            while (b >= s && S.id[S.perm[b]] > ida) {
                S.perm[b + 1] = S.perm[b];
                --b;
            }
            S.perm[b + 1] = pa;
        }
    }
}

// Gather every persistent per-particle array into sorted order (writes the
// other half of each double buffer).
__global__ __launch_bounds__(kBlock) void k_reorder(const DevScene* __restrict__ T, const Chunk* __restrict__ C) {
    const Chunk c = C[blockIdx.x];
    const DevScene& S = T[c.scene];
    const int s = c.begin + threadIdx.x;
    if (s >= S.n) return;
    const uint32_t o = S.perm[s];
    S.x2[s] = S.x[o];
    S.v2[s] = S.v[o];
    S.id2[s] = S.id[o];
    S.cell[s] = S.cellU[o];
    if (S.method == DFSPH) {
        S.k2[s] = S.k[o];
        S.kv2[s] = S.kv[o];
    }
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
    float sum = 0.0f;
    walk27(S.bCellStart, cb, S.nx, S.nx * S.ny, [&](int j) {
        const float4 y = S.xb[j];
        const float dx = xb.x - y.x, dy = xb.y - y.y, dz = xb.z - y.z;
        sum += W_cubic(dx * dx + dy * dy + dz * dz, S.kc);
    });
    bvol[b] = 1.0f / sum;
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
template <int M>
__global__ __launch_bounds__(kBlock) void k_build(const DevScene* __restrict__ T, const Chunk* __restrict__ C) {
    const Chunk c = C[blockIdx.x];
    const DevScene& S = T[c.scene];
    const int i = c.begin + threadIdx.x;
    if (i >= S.n) return;
    const KCoef k = S.kc;
    const float4 xi = S.x[i];
    const int ci = S.cell[i];
    const int nx = S.nx, nxny = S.nx * S.ny;
    uint32_t* L = list_of_w(S, i);
    const int maxN = S.maxNbr;
    int cnt = 0, ngr = 0;
    float ws = 0.f, sgx = 0.f, sgy = 0.f, sgz = 0.f, ssq = 0.f;
    walk27(S.cellStart, ci, nx, nxny, [&](int j) {
        if (j == i) return;
        const float4 xj = S.x[j];
        const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        if (!(r2 <= k.hhi2)) return;
        const float g = G_cubic(r2, k);
        bool gnz = (g != 0.0f);
        const float qf = sqrtf(r2) * k.invh;
        if (r2 >= k.hlo2 || fabsf(qf - 0.6666667f) < 1e-5f || r2 < 1e-9f) {
            const Exact e = exact_q(xi, xj, S.hD);
            if (!(e.q <= 1.0)) return;
            gnz = grad_nonzero64(e);
        }
        if (cnt < maxN) L[(size_t)cnt * kGroup] = (uint32_t)j;
        ++cnt;
        ngr += gnz ? 1 : 0;
        ws += W_cubic(r2, k);
        sgx += g * dx;
        sgy += g * dy;
        sgz += g * dz;
        ssq += g * g * r2;
    });
    float bw = 0.f, bgx = 0.f, bgy = 0.f, bgz = 0.f;
    float sx = 0.f, sy = 0.f, sz = 0.f;   // spiky boundary sum (PCISPH)
    float ax = 0.f, ay = 0.f, az = 0.f;   // adhesion sum
    const bool wantAdh = (S.flags & F_ADH) != 0;
    walk27(S.bCellStart, ci, nx, nxny, [&](int b) {
        const float4 y = S.xb[b];
        const float dx = xi.x - y.x, dy = xi.y - y.y, dz = xi.z - y.z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        if (!(r2 <= k.hhi2)) return;
        const float cb = S.rho0 * y.w;
        bw += cb * W_cubic(r2, k);
        const float g = G_cubic(r2, k);
        bgx += cb * g * dx;
        bgy += cb * g * dy;
        bgz += cb * g * dz;
        if (M == PCISPH) {
            const float gs = G_spiky(r2, k);
            sx += cb * gs * dx;
            sy += cb * gs * dy;
            sz += cb * gs * dz;
        }
        if (wantAdh && r2 > 0.0f) {
            const float rl = sqrtf(r2);
            const float s = cb * A_adhesion(rl, k) / rl;
            ax += s * dx;
            ay += s * dy;
            az += s * dz;
        }
    });
    if (cnt > maxN) atomicMax(&S.ss->overflow, (unsigned)cnt);
    S.nnb[i] = (uint16_t)min(cnt, maxN);
    S.ngrad[i] = (uint8_t)min(ngr, 255);
    const float rho = S.m * (S.W0 + ws) + bw;
    S.rho[i] = rho;
    S.gb[i] = make_float4(bgx, bgy, bgz, bw);
    if (wantAdh) S.adh[i] = make_float4(ax, ay, az, 0.f);
    if (M == DFSPH) {
        const float tx = S.m * sgx + bgx, ty = S.m * sgy + bgy, tz = S.m * sgz + bgz;
        const float den = tx * tx + ty * ty + tz * tz + S.m * S.m * ssq;
        S.alpha[i] = (fabsf(den) > 1e-5f) ? rho / den : 0.0f;
    } else if (M == WCSPH) {
        // p = max(0, kappa rho0/gamma ((rho/rho0)^gamma - 1)) with
        // (1+x)^g - 1 = expm1(g log1p(x)) and x = (rho - rho0)/rho0 exact-ish.
        const float x = (rho - S.rho0) / S.rho0;
        const float pr = fmaxf(0.0f, S.stiffness * S.rho0 / S.gammaEos * expm1f(S.gammaEos * log1pf(x)));
        S.p[i] = pr;
        S.xk[i] = make_float4(xi.x, xi.y, xi.z, pr / (rho * rho));
        const float4 vi = S.v[i];
        S.aux[i] = make_float4(vi.x, vi.y, vi.z, S.m / rho);
    } else {
        S.gbs[i] = make_float4(sx, sy, sz, 0.f);
        const float4 vi = S.v[i];
        S.aux[i] = make_float4(vi.x, vi.y, vi.z, S.m / rho);
    }
}

// ======================================================= non-pressure

// Akinci surface normals n_i = h sum_j (m/rho_j) gradW_ij
// (src/SurfaceTensionAkinci.h:46-61).  nrm.w carries rho_i for pass 2.
__global__ __launch_bounds__(kBlock) void k_normals(const DevScene* __restrict__ T, const Chunk* __restrict__ C) {
    const Chunk c = C[blockIdx.x];
    const DevScene& S = T[c.scene];
    const int i = c.begin + threadIdx.x;
    if (i >= S.n) return;
    const float4 xi = S.x[i];
    const uint32_t* L = list_of(S, i);
    const int cnt = S.nnb[i];
    float nx = 0.f, ny = 0.f, nz = 0.f;
    for (int q = 0; q < cnt; ++q) {
        const int j = (int)L[(size_t)q * kGroup];
        const float4 xj = S.x[j];
        const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
        const float g = G_cubic(dx * dx + dy * dy + dz * dz, S.kc) * (S.m / S.rho[j]);
        nx += g * dx;
        ny += g * dy;
        nz += g * dz;
    }
    const float h = S.kc.h;
    S.nrm[i] = make_float4(nx * h, ny * h, nz * h, S.rho[i]);
}

// Non-pressure accelerations (and, for WCSPH, the pressure force and the
// explicit integration).
//   pressure  src/WCSPHSolver.cpp:152-186
//   XSPH      src/XSPHViscosity.h:45-67
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
    if (i >= S.n) return;
    const KCoef k = S.kc;
    const float dt = (float)S.ss->dt;
    const float invdt = 1.0f / dt;
    const bool xs = (S.flags & F_XSPH) != 0, st = (S.flags & F_ST) != 0;
    const float4 xi = (M == WCSPH) ? S.xk[i] : S.x[i];
    const float4 vi = (M == WCSPH) ? S.aux[i] : S.v[i];
    const float rhoi = S.rho[i];
    const float4 ni = st ? S.nrm[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    const uint32_t* L = list_of(S, i);
    const int cnt = S.nnb[i];
    const float m = S.m;
    float px = 0.f, py = 0.f, pz = 0.f;   // pressure part
    float vx = 0.f, vy = 0.f, vz = 0.f;   // XSPH part (before nu/dt)
    float sx = 0.f, sy = 0.f, sz = 0.f;   // surface tension part (before gamma)
    for (int q = 0; q < cnt; ++q) {
        const int j = (int)L[(size_t)q * kGroup];
        const float4 xj = (M == WCSPH) ? S.xk[j] : S.x[j];
        const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        if (M == WCSPH) {
            const float g = G_cubic(r2, k) * m * (xi.w + xj.w);
            px -= g * dx;
            py -= g * dy;
            pz -= g * dz;
        }
        if (xs) {
            float4 vj;
            float mj;
            if (M == WCSPH) {
                vj = S.aux[j];
                mj = vj.w;
            } else {
                vj = S.v[j];
                mj = m / S.rho[j];
            }
            const float w = mj * W_cubic(r2, k);
            vx += w * (vj.x - vi.x);
            vy += w * (vj.y - vi.y);
            vz += w * (vj.z - vi.z);
        }
        if (st) {
            // list members satisfy |r| <= h exactly (FP64), as the
            // reference's test at src/SurfaceTensionAkinci.h:88
            const float rl = sqrtf(r2);
            if (r2 > 0.0f) {
                const float4 nj = S.nrm[j];
                const float kij = 2.0f * S.rho0 / (rhoi + nj.w);
                bool outer = 2.0f * rl > k.h;
                if (fabsf(2.0f * rl - k.h) < 1e-5f * k.h) outer = 2.0 * exact_q(xi, xj, S.hD).rl > S.hD;
                const float cs = -m * C_cohesion(rl, outer, k) / rl;
                sx += kij * (cs * dx - (ni.x - nj.x));
                sy += kij * (cs * dy - (ni.y - nj.y));
                sz += kij * (cs * dz - (ni.z - nj.z));
            }
        }
    }
    float ax = 0.f, ay = 0.f, az = 0.f;
    if (M == WCSPH) {
        const float4 gb = S.gb[i];
        ax = px - xi.w * gb.x;
        ay = py - xi.w * gb.y;
        az = pz - xi.w * gb.z;
    }
    if (xs) {
        const float f = S.nu * invdt;
        ax += f * vx;
        ay += f * vy;
        az += f * vz;
        if (S.nub != 0.0f) {
            const float fb = S.nub * (S.gb[i].w / rhoi) * invdt;
            ax -= fb * vi.x;
            ay -= fb * vi.y;
            az -= fb * vi.z;
        }
    }
    if (st) {
        ax += S.stCoef * sx;
        ay += S.stCoef * sy;
        az += S.stCoef * sz;
    }
    if (S.flags & F_ADH) {
        const float4 ad = S.adh[i];
        ax -= S.adhBeta * ad.x;
        ay -= S.adhBeta * ad.y;
        az -= S.adhBeta * ad.z;
    }
    S.a[i] = make_float4(ax, ay, az, 0.f);
    if (M == WCSPH) {
        const float nvx = vi.x + (ax + S.gx) * dt, nvy = vi.y + (ay + S.gy) * dt, nvz = vi.z + (az + S.gz) * dt;
        S.v[i] = make_float4(nvx, nvy, nvz, 0.f);
        S.x[i] = make_float4(xi.x + nvx * dt, xi.y + nvy * dt, xi.z + nvz * dt, 0.f);
    } else if (M == PCISPH) {
        const float wx = vi.x + (ax + S.gx) * dt, wy = vi.y + (ay + S.gy) * dt, wz = vi.z + (az + S.gz) * dt;
        S.xs[i] = make_float4(xi.x + wx * dt, xi.y + wy * dt, xi.z + wz * dt, 0.f);
    }
}

// WCSPH/PCISPH step-time bookkeeping: time += dt, steps++ (one thread per
// scene) and, for PCISPH, the entry test of the pressure loop
// (src/PCISPHSolver.cpp:365) with p = pacc = 0 (:83-120).
__global__ __launch_bounds__(kBlock) void k_step_begin(const DevScene* __restrict__ T, const Chunk* __restrict__ C, int* active) {
    const Chunk c = C[blockIdx.x];
    const DevScene& S = T[c.scene];
    const int i = c.begin + threadIdx.x;
    if (S.method == PCISPH && i < S.n) {
        S.p[i] = 0.0f;
        S.pacc[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    if (c.begin == 0 && threadIdx.x == 0) {
        Scalars* ss = S.ss;
        ss->steps += 1u;
        if (S.method == PCISPH) {
            ss->it = 0;
            const bool run = (ss->maxDensErr > S.eta || 0 < S.minIt) && 0 < S.maxIt;
            ss->done = run ? 0 : 1;
            if (run) atomicAdd(active, 1);
        }
    }
}

// ================================================================ DFSPH

// CFL time step (src/DFSPHSolver.cpp:827-870): maxVel = max(0.1, max_i
// sqrt(|v_i + (a_i + g) dt|)); dt' = cfl if cfl < dt, else 0.99 dt after
// more than 3 density iterations, else 1.001 dt; clamped to [min, max].
__global__ __launch_bounds__(kBlock) void k_dt(const DevScene* __restrict__ T, const Chunk* __restrict__ C) {
    const Chunk c = C[blockIdx.x];
    const DevScene& S = T[c.scene];
    Scalars* ss = S.ss;
    __shared__ double sh[4];
    const int i = c.begin + threadIdx.x;
    const float dt = (float)ss->dt;
    double val = 0.0;
    if (i < S.n) {
        const float4 v = S.v[i];
        float ax = S.gx, ay = S.gy, az = S.gz;
        if (S.flags & (F_XSPH | F_ST | F_ADH)) {
            const float4 a = S.a[i];
            ax += a.x;
            ay += a.y;
            az += a.z;
        }
        const float cx = v.x + ax * dt, cy = v.y + ay * dt, cz = v.z + az * dt;
        val = (double)sqrtf(sqrtf(cx * cx + cy * cy + cz * cz));
    }
    const double bm = block_max(val, sh);
    if (threadIdx.x == 0) S.partial[c.begin / kBlock] = bm;
    if (arrive_last(&ss->ctr, (unsigned)S.nChunks)) {
        const double mv = scene_partial_max(S.partial, S.nChunks, sh);
        if (threadIdx.x == 0) {
            const double maxVel = fmax(0.1, mv);
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
            ss->ctr = 0u;
            ss->steps += 1u;
        }
    }
}

// predictVelocities (src/DFSPHSolver.cpp:191-214) fused with the first loop
// of densityWarmStart (:602-613): k_i <- min(k_i/rho_i, 1e-9)/dt^2, staged
// in xk.w for the warm-start sweep.
__global__ __launch_bounds__(kBlock) void k_df_predict(const DevScene* __restrict__ T, const Chunk* __restrict__ C) {
    const Chunk c = C[blockIdx.x];
    const DevScene& S = T[c.scene];
    const int i = c.begin + threadIdx.x;
    if (i >= S.n) return;
    const float dt = (float)S.ss->dt;
    float4 v = S.v[i];
    float ax = S.gx, ay = S.gy, az = S.gz;
    if (S.flags & (F_XSPH | F_ST | F_ADH)) {
        const float4 a = S.a[i];
        ax = a.x + S.gx;
        ay = a.y + S.gy;
        az = a.z + S.gz;
    }
    v.x += ax * dt;
    v.y += ay * dt;
    v.z += az * dt;
    S.v[i] = v;
    const float kw = -fmaxf(-S.k[i] / S.rho[i], -1.0e-9f) / (dt * dt);
    const float4 x = S.x[i];
    S.xk[i] = make_float4(x.x, x.y, x.z, kw);
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
    const KCoef k = S.kc;
    const float eps = 1.0e-5f;
    const float dt = (float)ss->dt;
    const float4 xi = S.xk[i];
    const float si = xi.w;
    const uint32_t* L = list_of(S, i);
    const int cnt = S.nnb[i];
    float sx = 0.f, sy = 0.f, sz = 0.f;
    for (int q = 0; q < cnt; ++q) {
        const int j = (int)L[(size_t)q * kGroup];
        const float4 xj = S.xk[j];
        const float ks = si + xj.w;
        if (fabsf(ks) > eps) {
            const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
            const float g = S.m * ks * G_cubic(dx * dx + dy * dy + dz * dz, k);
            sx += g * dx;
            sy += g * dy;
            sz += g * dz;
        }
    }
    const float cb = (MODE == 0) ? 2.0f * si : si;
    if (fabsf(cb) > eps) {
        const float4 gb = S.gb[i];
        sx += cb * gb.x;
        sy += cb * gb.y;
        sz += cb * gb.z;
    }
    float4 v = S.v[i];
    v.x -= dt * sx;
    v.y -= dt * sy;
    v.z -= dt * sz;
    S.v[i] = v;
    if (MODE == 0) S.k[i] = 0.0f;
    if (MODE == 1) S.k[i] += (S.rhoS[i] - S.rho0) * S.alpha[i];
    if (MODE == 2) S.kv[i] += S.divErr[i] * S.alpha[i];
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
__global__ __launch_bounds__(kBlock) void k_df_err(const DevScene* __restrict__ T, const Chunk* __restrict__ C, int* active) {
    const Chunk c = C[blockIdx.x];
    const DevScene& S = T[c.scene];
    Scalars* ss = S.ss;
    constexpr bool DENS = (MODE <= 1);
    if (MODE == 1 && ss->done) return;
    if (MODE == 3 && ss->doneV) return;
    __shared__ double sh[4];
    const int i = c.begin + threadIdx.x;
    const KCoef k = S.kc;
    const float dt = (float)ss->dt;
    double e = 0.0;
    if (i < S.n) {
        const float4 xi = S.x[i];
        const float4 vi = S.v[i];
        const uint32_t* L = list_of(S, i);
        const int cnt = S.nnb[i];
        float dc = 0.f;
        for (int q = 0; q < cnt; ++q) {
            const int j = (int)L[(size_t)q * kGroup];
            const float4 xj = S.x[j];
            const float4 vj = S.v[j];
            const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
            const float g = G_cubic(dx * dx + dy * dy + dz * dz, k);
            dc += g * ((vi.x - vj.x) * dx + (vi.y - vj.y) * dy + (vi.z - vj.z) * dz);
        }
        const float4 gb = S.gb[i];
        dc = S.m * dc + (vi.x * gb.x + vi.y * gb.y + vi.z * gb.z);
        const float rho = S.rho[i], al = S.alpha[i];
        float s;
        if (DENS) {
            const float rs = fmaxf(S.rho[i] + dc * dt, S.rho0);
            S.rhoS[i] = rs;
            s = (rs - S.rho0) * al / (dt * dt * rho);
            e = ((double)rs - (double)S.rho0) / (double)S.rho0;
        } else {
            const float de = (S.ngrad[i] < 20) ? 0.0f : fmaxf(dc, 0.0f);
            S.divErr[i] = de;
            s = de * al / (dt * rho);
            e = (double)de / (double)S.rho0;
        }
        S.xk[i] = make_float4(xi.x, xi.y, xi.z, s);
    }
    const double bs = block_sum(e, sh);
    if (threadIdx.x == 0) S.partial[c.begin / kBlock] = bs;
    if (arrive_last(&ss->ctr, (unsigned)S.nChunks)) {
        const double tot = scene_partial_sum(S.partial, S.nChunks, sh);
        if (threadIdx.x == 0) {
            const double avg = (S.n == 0) ? 0.0 : tot / (double)S.n;
            ss->ctr = 0u;
            if (DENS) {
                int it;
                double err;
                if (MODE == 0) {
                    ss->avgErr0 = avg;
                    ss->it = 0;
                    it = 0;
                    err = 0.0;
                } else {
                    it = ++ss->it;
                    ss->avgErr = avg;
                    err = avg;
                }
                const bool more = ((err > S.eta) || (it < S.minIt)) && (it < S.maxIt);
                ss->done = more ? 0 : 1;
                if (MODE == 0 && more) atomicAdd(active, 1);
                if (MODE == 1 && !more) atomicSub(active, 1);
                if (!more) ss->sumIt += (unsigned)it;
            } else {
                int it;
                double err;
                if (MODE == 2) {
                    ss->avgErrV0 = avg;
                    ss->itV = 0;
                    it = 0;
                    err = 0.0;
                    ss->time += ss->dt;
                } else {
                    it = ++ss->itV;
                    ss->avgErrV = avg;
                    err = avg;
                }
                const bool more = ((err > S.etaV / ss->dt) || (it < S.minItV)) && (it < S.maxIt);
                ss->doneV = more ? 0 : 1;
                if (MODE == 2 && more) atomicAdd(active, 1);
                if (MODE == 3 && !more) atomicSub(active, 1);
                if (!more) ss->sumItV += (unsigned)it;
            }
        }
    }
}

// updatePositions (src/DFSPHSolver.cpp:216-236) fused with the cell key and
// histogram of the grid rebuild.
__global__ __launch_bounds__(kBlock) void k_df_positions(const DevScene* __restrict__ T, const Chunk* __restrict__ C) {
    const Chunk c = C[blockIdx.x];
    const DevScene& S = T[c.scene];
    const int i = c.begin + threadIdx.x;
    if (i >= S.n) return;
    const float dt = (float)S.ss->dt;
    float4 x = S.x[i];
    const float4 v = S.v[i];
    x.x += v.x * dt;
    x.y += v.y * dt;
    x.z += v.z * dt;
    S.x[i] = x;
    const int cl = cell_of(S, x);
    S.cellU[i] = cl;
    atomicAdd(&S.cellCount[cl], 1);
}

// =============================================================== PCISPH

// Predicted densities at x* with the candidate cells of x
// (src/PCISPHSolver.cpp:164-212; every candidate of the 27 cells, self
// included), then updatePressure (:249-295): p_i += delta max(0, rho_i -
// rho0); the scene's max relative error decides the loop (:365), counting
// this iteration as done (its pressure acceleration always follows).
__global__ __launch_bounds__(kBlock) void k_pci_density(const DevScene* __restrict__ T, const Chunk* __restrict__ C, int* active) {
    const Chunk c = C[blockIdx.x];
    const DevScene& S = T[c.scene];
    Scalars* ss = S.ss;
    if (ss->done) return;
    __shared__ double sh[4];
    const int i = c.begin + threadIdx.x;
    const KCoef k = S.kc;
    double e = 0.0;
    if (i < S.n) {
        const float4 pi = S.xs[i];
        const int ci = S.cell[i];
        const int nx = S.nx, nxny = S.nx * S.ny;
        float ws = 0.f;
        walk27(S.cellStart, ci, nx, nxny, [&](int j) {
            const float4 pj = S.xs[j];
            const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            ws += W_cubic(dx * dx + dy * dy + dz * dz, k);
        });
        float bw = 0.f;
        walk27(S.bCellStart, ci, nx, nxny, [&](int b) {
            const float4 y = S.xb[b];
            const float dx = pi.x - y.x, dy = pi.y - y.y, dz = pi.z - y.z;
            bw += y.w * W_cubic(dx * dx + dy * dy + dz * dz, k);
        });
        const float rho = S.m * ws + S.rho0 * bw;
        S.rho[i] = rho;
        const float err = fmaxf(0.0f, rho - S.rho0);
        const float pr = S.p[i] + S.pciDelta * err;
        S.p[i] = pr;
        const float4 xi = S.x[i];
        S.xk[i] = make_float4(xi.x, xi.y, xi.z, pr / (rho * rho));
        e = (double)err / (double)S.rho0;
    }
    const double bm = block_max(e, sh);
    if (threadIdx.x == 0) S.partial[c.begin / kBlock] = bm;
    if (arrive_last(&ss->ctr, (unsigned)S.nChunks)) {
        const double mx = scene_partial_max(S.partial, S.nChunks, sh);
        if (threadIdx.x == 0) {
            ss->ctr = 0u;
            ss->maxDensErr = mx;
            const int it = ++ss->it;
            const bool more = ((mx > S.eta) || (it < S.minIt)) && (it < S.maxIt);
            if (!more) {
                ss->done = 1;
                ss->sumIt += (unsigned)it;
                atomicSub(active, 1);
            }
        }
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
    const KCoef k = S.kc;
    const float dt = (float)ss->dt;
    const float4 xi = S.xk[i];
    const uint32_t* L = list_of(S, i);
    const int cnt = S.nnb[i];
    float ax = 0.f, ay = 0.f, az = 0.f;
    for (int q = 0; q < cnt; ++q) {
        const int j = (int)L[(size_t)q * kGroup];
        const float4 xj = S.xk[j];
        const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
        const float g = S.m * (xi.w + xj.w) * G_spiky(dx * dx + dy * dy + dz * dz, k);
        ax -= g * dx;
        ay -= g * dy;
        az -= g * dz;
    }
    const float4 gs = S.gbs[i];
    ax -= xi.w * gs.x;
    ay -= xi.w * gs.y;
    az -= xi.w * gs.z;
    S.pacc[i] = make_float4(ax, ay, az, 0.f);
    const float4 a = S.a[i];
    const float4 v = S.v[i];
    const float wx = v.x + ((a.x + ax) + S.gx) * dt;
    const float wy = v.y + ((a.y + ay) + S.gy) * dt;
    const float wz = v.z + ((a.z + az) + S.gz) * dt;
    S.xs[i] = make_float4(xi.x + wx * dt, xi.y + wy * dt, xi.z + wz * dt, 0.f);
}

// PCISPH integrate (src/PCISPHSolver.cpp:444-474): a += g + pacc.
__global__ __launch_bounds__(kBlock) void k_pci_integrate(const DevScene* __restrict__ T, const Chunk* __restrict__ C) {
    const Chunk c = C[blockIdx.x];
    const DevScene& S = T[c.scene];
    const int i = c.begin + threadIdx.x;
    if (i >= S.n) return;
    const float dt = (float)S.ss->dt;
    const float4 a = S.a[i], pa = S.pacc[i];
    float4 v = S.v[i];
    float4 x = S.x[i];
    v.x += (a.x + (S.gx + pa.x)) * dt;
    v.y += (a.y + (S.gy + pa.y)) * dt;
    v.z += (a.z + (S.gz + pa.z)) * dt;
    x.x += v.x * dt;
    x.y += v.y * dt;
    x.z += v.z * dt;
    S.v[i] = v;
    S.x[i] = x;
    if (c.begin == 0 && threadIdx.x == 0) S.ss->time += S.ss->dt;
}

// WCSPH: time += dt after the integration (src/WCSPHSolver.cpp:77).
__global__ void k_time_advance(const DevScene* __restrict__ T, int nscenes) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nscenes) T[s].ss->time += T[s].ss->dt;
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
__global__ void k_unsort_f1(const float* __restrict__ src, const uint32_t* __restrict__ id, float* __restrict__ dst, int n) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    dst[id[s]] = src[s];
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

void launch_keys(hipStream_t st, const DevScene* T, const Chunk* C, int n) { if (n) LAUNCH(k_keys, n, st, T, C); }
void launch_scan(hipStream_t st, const DevScene* T, const Chunk* CC, int ncc, int nscenes) {
    if (!ncc) return;
    LAUNCH(k_scan_sum, ncc, st, T, CC);
    LAUNCH(k_scan_offsets, nscenes, st, T, nscenes);
    LAUNCH(k_scan_final, ncc, st, T, CC);
}
void launch_scatter(hipStream_t st, const DevScene* T, const Chunk* C, int n) { if (n) LAUNCH(k_scatter, n, st, T, C); }
void launch_cell_sort(hipStream_t st, const DevScene* T, const Chunk* CC, int ncc) { if (ncc) LAUNCH(k_cell_sort, ncc, st, T, CC); }
void launch_reorder(hipStream_t st, const DevScene* T, const Chunk* C, int n) { if (n) LAUNCH(k_reorder, n, st, T, C); }
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
void launch_step_begin(hipStream_t st, const DevScene* T, const Chunk* C, int n, int* active) { if (n) LAUNCH(k_step_begin, n, st, T, C, active); }
void launch_dt(hipStream_t st, const DevScene* T, const Chunk* C, int n) { if (n) LAUNCH(k_dt, n, st, T, C); }
void launch_df_predict(hipStream_t st, const DevScene* T, const Chunk* C, int n) { if (n) LAUNCH(k_df_predict, n, st, T, C); }
void launch_df_vel(hipStream_t st, int mode, const DevScene* T, const Chunk* C, int n) {
    if (!n) return;
    if (mode == 0) LAUNCH(k_df_vel<0>, n, st, T, C);
    else if (mode == 1) LAUNCH(k_df_vel<1>, n, st, T, C);
    else LAUNCH(k_df_vel<2>, n, st, T, C);
}
void launch_df_err(hipStream_t st, int mode, const DevScene* T, const Chunk* C, int n, int* active) {
    if (!n) return;
    if (mode == 0) LAUNCH(k_df_err<0>, n, st, T, C, active);
    else if (mode == 1) LAUNCH(k_df_err<1>, n, st, T, C, active);
    else if (mode == 2) LAUNCH(k_df_err<2>, n, st, T, C, active);
    else LAUNCH(k_df_err<3>, n, st, T, C, active);
}
void launch_df_positions(hipStream_t st, const DevScene* T, const Chunk* C, int n) { if (n) LAUNCH(k_df_positions, n, st, T, C); }
void launch_pci_density(hipStream_t st, const DevScene* T, const Chunk* C, int n, int* active) { if (n) LAUNCH(k_pci_density, n, st, T, C, active); }
void launch_pci_pacc(hipStream_t st, const DevScene* T, const Chunk* C, int n, int iter) { if (n) LAUNCH(k_pci_pacc, n, st, T, C, iter); }
void launch_pci_integrate(hipStream_t st, const DevScene* T, const Chunk* C, int n) { if (n) LAUNCH(k_pci_integrate, n, st, T, C); }
void launch_time_advance(hipStream_t st, const DevScene* T, int nscenes) {
    hipLaunchKernelGGL(k_time_advance, dim3((nscenes + 255) / 256), dim3(256), 0, st, T, nscenes);
}
void launch_fill_int(hipStream_t st, int* p, int v, long n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_int, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, v, n);
}
void launch_unsort_f4(hipStream_t st, const float4* src, const uint32_t* id, float* dst, int n, int comps) {
    if (n) hipLaunchKernelGGL(k_unsort_f4, dim3((n + 255) / 256), dim3(256), 0, st, src, id, dst, n, comps);
}
void launch_unsort_f1(hipStream_t st, const float* src, const uint32_t* id, float* dst, int n) {
    if (n) hipLaunchKernelGGL(k_unsort_f1, dim3((n + 255) / 256), dim3(256), 0, st, src, id, dst, n);
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
