// sphg_api.hip -- C ABI of libsphg: contexts, scenes, batched steps.
//
// The step sequences restate the reference solvers' phase order:
//   WCSPH  src/WCSPHSolver.cpp:14-78
//   PCISPH src/PCISPHSolver.cpp:21-81 (loop :360-387)
//   DFSPH  src/DFSPHSolver.cpp:38-112 (loops :311-406, :494-590)
// The iteration loops are driven from the host here, reading one int (the
// number of scenes still iterating) per iteration.

#include "../../include/sphg.h"
#include "sphg_internal.h"
#include "sphg_kernels.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <mutex>
#include <string>
#include <vector>

using namespace sphg;

// Device memory of the allocation-free batch step (sphg_batch_step_noalloc):
// the batch's two scene tables and its particle and cell chunk tables live
// here instead of in memory the call allocates.  Each device that the
// library runs on holds one copy until the process ends; host code uses its
// address from hipGetSymbolAddress (scratch_address), never this name.
static __device__ __align__(256) unsigned char scratch[SPHG_NOALLOC_SCRATCH_BYTES];
// the table sizes sphg.h states for it
static_assert(sizeof(sphg::DevScene) == 928 && sizeof(sphg::Chunk) == 8, "update the sizes in sphg.h");

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail(SPHG_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));          \
    } while (0)

constexpr int kDefaultMaxNbr = 96;

// ----------------------------------------------------------- profiling

enum Kind {
    K_KEYS, K_SCAN, K_SCATTER, K_FINALIZE, K_REORDER, K_BUILD, K_NORMALS, K_FORCES, K_DT,
    K_DFPREDICT, K_DFVEL_WARM, K_DFVEL_DENS, K_DFVEL_DIV, K_DFERR_DENS, K_DFERR_DIV, K_DFPOS, K_PCIDENS,
    K_PCIPACC, K_PCIINT, K_TIME, K_CUBE, K_EMIT, K_PB, K_NKINDS
};
const char* kKindNames[K_NKINDS] = {
    "keys", "scan", "scatter", "finalize", "rank_reorder", "build", "normals", "forces", "dt",
    "df_predict", "df_vel_warm", "df_vel_density", "df_vel_divergence", "df_err_density", "df_err_divergence",
    "df_positions", "pci_density", "pci_pacc", "pci_integrate", "time_advance", "cube", "emit", "pcisph_boundary"};

struct Profiler {
    bool on = false;
    uint64_t mask = ~0ull;      // kinds recorded while on
    uint32_t stride = 1;        // record every stride-th launch of each kind
    uint64_t seen[K_NKINDS] = {};   // launches of each kind since enabling
    struct Rec {
        int kind;
        hipEvent_t a, b;
    };
    std::vector<Rec> pending;
    std::vector<hipEvent_t> pool;
    uint64_t count[K_NKINDS] = {};
    double ms[K_NKINDS] = {};
    hipEvent_t get() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        // no system-scope fence on record: a timing event then adds no
        // cache writeback/invalidate between the kernels it brackets
        hipEvent_t e;
        (void)hipEventCreateWithFlags(&e, hipEventDisableSystemFence);
        return e;
    }
    void collect() {
        if (pending.empty()) return;
        for (auto& r : pending) {
            (void)hipEventSynchronize(r.b);
            float t = 0.f;
            (void)hipEventElapsedTime(&t, r.a, r.b);
            count[r.kind] += 1;
            ms[r.kind] += t;
            pool.push_back(r.a);
            pool.push_back(r.b);
        }
        pending.clear();
    }
    ~Profiler() {
        for (auto& r : pending) {
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

// Device memory of scenes, batch tables and staging is stream-ordered
// (hipMallocAsync / hipFreeAsync on the owning context's stream): a plain
// hipMalloc or hipFree waits for the work queued on every stream of the
// device, so admitting or retiring a scene would stall the other contexts'
// scenes (measured: 2-6 ms per call while another context steps a 1M scene).
// Freed blocks stay in the device's default pool (release threshold raised
// at context creation) for the next allocation.
hipError_t dev_alloc(void** p, size_t bytes, hipStream_t st) { return hipMallocAsync(p, bytes, st); }
void dev_free(void* p, hipStream_t st) {
    if (p) (void)hipFreeAsync(p, st);
}

// A temporary device allocation freed on every return path.
struct DevTemp {
    void* p = nullptr;
    hipStream_t st;
    explicit DevTemp(hipStream_t s) : st(s) {}
    DevTemp(const DevTemp&) = delete;
    DevTemp& operator=(const DevTemp&) = delete;
    hipError_t alloc(size_t bytes) { return dev_alloc(&p, bytes, st); }
    ~DevTemp() { dev_free(p, st); }
    float* f() const { return (float*)p; }
};

struct Scope {
    Profiler* p;
    hipStream_t st;
    int kind;
    hipEvent_t a{}, b{};
    bool rec;
    Scope(Profiler* p_, hipStream_t s, int k)
        : p(p_), st(s), kind(k), rec(p_->on && ((p_->mask >> k) & 1) && (p_->seen[k]++ % p_->stride == 0)) {
        if (rec) {
            a = p->get();
            (void)hipEventRecord(a, st);
        }
    }
    ~Scope() {
        if (rec) {
            b = p->get();
            (void)hipEventRecord(b, st);
            p->pending.push_back({kind, a, b});
            if (p->pending.size() > 4096) p->collect();
        }
    }
};

}  // namespace

struct sphg_ctx {
    int device = 0;
    unsigned char* scratch = nullptr;   // this device's address of `scratch`
    hipStream_t stream = nullptr;
    // snapshot read-backs (default priority); made with the context's first
    // snapshot, so a context that takes none holds one hardware queue
    hipStream_t copyStream = nullptr;
    bool graphs = false;                // replay captured step segments (persistent batches)
    int* dActive = nullptr;   // [0] scenes still iterating, [1] largest iteration count reached
    int* hActive = nullptr;   // pinned copy
    Profiler prof;
    // the context's scenes; creating and destroying a scene that no running
    // call uses may happen on another thread than the one running the
    // context, so the list has a lock of its own
    std::mutex scenesMu;
    std::vector<sphg_scene*> scenes;
};

namespace {

struct Arena {
    char* base = nullptr;
    size_t size = 0, used = 0;
    size_t reserve(size_t bytes) {
        const size_t off = (used + 255) & ~size_t(255);
        used = off + bytes;
        return off;
    }
};

struct Batch {
    std::vector<sphg_scene*> scenes;
    DevScene* dT[2] = {nullptr, nullptr};
    Chunk* dP = nullptr;
    Chunk* dC = nullptr;
    int nP = 0, nC = 0;
    int maxCellChunks = 0;     // largest cell-chunk count of one scene
    int maxIt = 0;
    int minIt = 0, minItV = 0;   // smallest minimum iteration counts in the batch
    int lastIt = 0, lastItV = 0; // iterations the loops needed in the previous step
    std::vector<int> parity;   // parity of each scene when the tables were built
    std::vector<uint32_t> nAt; // active count of each scene in the tables
    // Fixed launch sequences of the step, captured once per table parity
    // (which of the two tables is dT[0]) and replayed; only for batches that
    // live across calls
    bool graphs = false;
    DevScene* tab0 = nullptr;
    // the tables are the library's allocations (false: they point into the
    // allocation-free step's scratch array, which is not freed)
    bool ownTables = true;
    // launch schedule (any choice gives the same results): loop iterations
    // launched past the previous step's count before the first read-back,
    // and segment replay (-1: the context's setting)
    int slackD = 1, slackV = 0, replay = -1;
    hipGraphExec_t seg[3][2] = {};
    hipStream_t allocStream = nullptr;  // the stream the tables were allocated on
    ~Batch() {
        for (auto& a : seg)
            for (auto& e : a)
                if (e) (void)hipGraphExecDestroy(e);
        if (!ownTables) return;
        dev_free(dT[0], allocStream);
        dev_free(dT[1], allocStream);
        dev_free(dP, allocStream);
        dev_free(dC, allocStream);
    }
};

}  // namespace

struct sphg_scene {
    sphg_ctx* ctx = nullptr;
    sphg_scene_desc desc{};
    DevScene view[2];          // view[p]: buffers of parity p current
    int par = 0;
    bool ready = false;
    // DFSPH: the last call stepped the scene, so the fields a step forms
    // before its grid rebuild (a, normals, rho*) are in the particle order
    // of the other half's ids (the order before the rebuild's sort)
    bool stepped = false;
    std::vector<float> fxyz, fvel, bxyz;
    std::vector<uint32_t> border;   // sorted boundary slot -> input index
    std::vector<sphg_emitter_desc> emitters;
    std::vector<float> bnrm;        // PCISPH boundary method: one normal per boundary particle
    uint32_t n = 0, nb = 0;         // n: active particles
    uint32_t cap = 0;               // particle slots: fluid + every emitter's budget
    double dt0 = 0, dtMin = 0, dtMax = 0;
    char* mem = nullptr;            // stream-ordered on ctx->stream
    size_t memBytes = 0;
    hipStream_t memStream = nullptr;
    float* tmp = nullptr;      // 4 floats per particle of staging
    Scalars* ss = nullptr;
    Batch* self = nullptr;
    std::vector<sphg_snapshot*> snaps;
    ~sphg_scene();
};

struct sphg_snapshot {
    sphg_scene* scene = nullptr;
    float* dStage = nullptr;       // x[cap*3], v[cap*3], then a Scalars copy
    float* hStage = nullptr;       // pinned mirror of dStage
    size_t capBytes = 0;           // bytes of one of x, v
    uint32_t n = 0;
    bool pending = false;
    hipEvent_t ready = nullptr, done = nullptr;
    void release() {
        if (done) (void)hipEventSynchronize(done);
        if (dStage) (void)hipFree(dStage);
        if (hStage) (void)hipHostFree(hStage);
        if (ready) (void)hipEventDestroy(ready);
        if (done) (void)hipEventDestroy(done);
        dStage = hStage = nullptr;
        ready = done = nullptr;
    }
};

sphg_scene::~sphg_scene() {
    for (auto* sn : snaps) {
        sn->release();
        sn->scene = nullptr;
    }
    delete self;
    dev_free(mem, memStream);
}

namespace {

// PCISPH scaling factor delta (src/PCISPHSolver.cpp:389-421): Poly6
// gradients over a 3x3x3 prototype at spacing 2r.
double pcisph_delta(double r, double h, double m, double rho0, double dt) {
    const double h2 = h * h;
    const double h9 = h * h * h * h * h * h * h * h * h;
    const double gc = 945.0 / (32.0 * M_PI * h9);
    const double dist = 2.0 * r;
    double sx = 0, sy = 0, sz = 0, sq = 0;
    for (int i = -1; i <= 1; i++)
        for (int j = -1; j <= 1; j++)
            for (int k = -1; k <= 1; k++) {
                const double rx = -i * dist, ry = -j * dist, rz = -k * dist;
                const double ml = std::sqrt(rx * rx + ry * ry + rz * rz);
                const double ml2 = ml * ml;
                double gx = 0, gy = 0, gz = 0;
                if (ml2 <= h2) {
                    const double d = h2 - ml2;
                    const double f = -d * d * gc;
                    gx = f * rx;
                    gy = f * ry;
                    gz = f * rz;
                }
                sx += gx;
                sy += gy;
                sz += gz;
                sq += gx * gx + gy * gy + gz * gz;
            }
    const double beta = 2 * std::pow(m * dt / rho0, 2);
    const double dot = sx * sx + sy * sy + sz * sz;
    return -1.0 / (beta * (-dot - sq));
}

int check_desc(const sphg_scene_desc& d) {
    if (d.method < 0 || d.method > 2) return fail(SPHG_EINVAL, "simulationMethod must be 0 (WCSPH), 1 (PCISPH) or 2 (DFSPH)");
    if (d.boundary_method != SPHG_BOUNDARY_AKINCI && d.boundary_method != SPHG_BOUNDARY_CUBE &&
        d.boundary_method != SPHG_BOUNDARY_PCISPH)
        return fail(SPHG_EINVAL, "boundaryMethod must be 0 (cube), 1 (PCISPH) or 2 (Akinci)");
    if (d.boundary_method == SPHG_BOUNDARY_PCISPH && d.adhesion_method == SPHG_ADHESION_AKINCI && d.adhesion != 0.0)
        return fail(SPHG_EINVAL, "adhesion needs Akinci boundary volumes; with the PCISPH boundary method the reference "
                                 "reads them from a model that has none (src/AdhesionForce.h:46-50)");
    if (!std::isfinite(d.pb_normal_fct) || !std::isfinite(d.pb_tangential_fct))
        return fail(SPHG_EINVAL, "the PCISPH boundary normal and tangential factors must be finite");
    if (d.cube_enabled && d.boundary_method != SPHG_BOUNDARY_CUBE)
        return fail(SPHG_EINVAL, "a cube clamp needs boundaryMethod 0 (cube)");
    if (d.cube_enabled)
        for (int a = 0; a < 3; ++a)
            if (!(d.cube_min[a] <= d.cube_max[a])) return fail(SPHG_EINVAL, "cube_min must not exceed cube_max");
    if (!(d.particle_radius > 0)) return fail(SPHG_EINVAL, "particleRadius must be > 0");
    if (!(d.density0 > 0)) return fail(SPHG_EINVAL, "density0 must be > 0");
    if (!(d.time_step > 0)) return fail(SPHG_EINVAL, "timeStep must be > 0");
    if (!(d.fps > 0)) return fail(SPHG_EINVAL, "fps must be > 0");
    if (d.method == 0 && !(d.gamma != 0)) return fail(SPHG_EINVAL, "WCSPH gamma must be nonzero");
    if (d.viscosity_method > SPHG_VISCOSITY_XSPH && d.viscosity != 0.0)
        return fail(SPHG_EINVAL, "viscosityMethod must be -1, 0 (standard), 1 (artificial) or 2 (XSPH)");
    if (d.max_neighbors > 65535) return fail(SPHG_EINVAL, "max_neighbors must be <= 65535");
    return SPHG_OK;
}

void fill_coef_d(KCoefD& k, double h) {
    const double pi = M_PI;
    const double h3 = h * h * h;
    const double cl = 48.0 / (pi * h3);
    const double h6 = h3 * h3;
    const double h9 = h6 * h3;
    k.h = h;
    k.invh = 1.0 / h;
    k.h2 = h * h;
    k.ck = 8.0 / (pi * h3);
    k.clh2 = cl / (h * h);
    k.clh = cl / h;
    k.spg = 45.0 / (pi * h6);
    k.co1 = 32.0 / (pi * h9);
    k.co2 = h6 / 64.0;
    k.hlo2 = h * h * (1.0 - 1e-9);
    k.hhi2 = h * h * (1.0 + 1e-9);
    k.gA = 3.0 * k.clh2 / h;
    k.gB = -2.0 * k.clh2;
    k.hHalf = 0.5 * h;
    k.vl = 45.0 / (M_PI * h * h * h * h * h * h);
    k.p6v = 315.0 / (64.0 * M_PI * h9);
}

void fill_coef(KCoef& k, double h) {
    const double pi = M_PI;
    const double h3 = h * h * h;
    const double cl = 48.0 / (pi * h3);
    const double h6 = h3 * h3;
    const double h9 = h6 * h3;
    k.h = (float)h;
    k.invh = (float)(1.0 / h);
    k.h2 = (float)(h * h);
    k.ck = (float)(8.0 / (pi * h3));
    k.clh2 = (float)(cl / (h * h));
    k.clh = (float)(cl / h);
    k.spg = (float)(45.0 / (pi * h6));
    k.co1 = (float)(32.0 / (pi * h9));
    k.co2 = (float)(h6 / 64.0);
    k.hlo2 = (float)(h * h * (1.0 - 1e-4));
    k.hhi2 = (float)(h * h * (1.0 + 1e-4));
}

constexpr int kXcds = 8;

std::vector<Chunk> xcd_order(const std::vector<Chunk>& in) {
    const size_t n = in.size();
    std::vector<Chunk> out(n);
    const size_t per = (n + kXcds - 1) / kXcds;   // chunks per XCD (the last XCD may get fewer)
    size_t pos = 0;
    // launch slot L runs on XCD L % 8; fill XCD x's slots with chunks
    // [x * per, (x + 1) * per) in order
    std::vector<size_t> next(kXcds);
    for (int x = 0; x < kXcds; ++x) next[x] = x * per;
    for (size_t L = 0; L < n; ++L) {
        int x = (int)(L % kXcds);
        // an XCD whose range is exhausted hands its slot to the next one with work left
        int tries = 0;
        while (next[x] >= std::min(n, (size_t)(x + 1) * per) && tries < kXcds) {
            x = (x + 1) % kXcds;
            ++tries;
        }
        out[pos++] = in[next[x]++];
    }
    return out;
}

int upload_batch_tables(sphg_ctx* ctx, Batch& B) {
    const int ns = (int)B.scenes.size();
    std::vector<DevScene> t0(ns), t1(ns);
    std::vector<Chunk> P, Cc;
    B.parity.resize(ns);
    B.nAt.resize(ns);
    B.maxIt = 0;
    B.minIt = 1 << 30;
    B.minItV = 1 << 30;
    int maxCC = 0;
    for (int s = 0; s < ns; ++s) {
        sphg_scene* sc = B.scenes[s];
        t0[s] = sc->view[sc->par];
        t1[s] = sc->view[1 - sc->par];
        B.parity[s] = sc->par;
        B.nAt[s] = sc->n;
        for (int c = 0; c < t0[s].nChunks; ++c) P.push_back(Chunk{s, c * kBlock});
        for (int c = 0; c < t0[s].nCellChunks; ++c) Cc.push_back(Chunk{s, c * kCellChunk});
        maxCC = std::max(maxCC, t0[s].nCellChunks);
        B.maxIt = std::max(B.maxIt, std::max(t0[s].maxIt, std::max(t0[s].maxItV, 0)));
        B.minIt = std::min(B.minIt, std::min(t0[s].minIt, t0[s].maxIt));
        B.minItV = std::min(B.minItV, std::min(t0[s].minItV, t0[s].maxIt));
    }
    // Blocks are dispatched round-robin over the XCDs (block b -> XCD b % 8),
    // each with its own L2: order the table so that XCD k runs one
    // contiguous, spatially coherent range of chunks and finds its
    // neighbors' data in its own L2.
    P = xcd_order(P);
    Cc = xcd_order(Cc);
    if (!B.dT[0]) {
        B.allocStream = ctx->stream;
        HIP_TRY(dev_alloc((void**)&B.dT[0], sizeof(DevScene) * ns, ctx->stream));
        HIP_TRY(dev_alloc((void**)&B.dT[1], sizeof(DevScene) * ns, ctx->stream));
        B.tab0 = B.dT[0];
    }
    if (!B.dP) {
        B.allocStream = ctx->stream;
        HIP_TRY(dev_alloc((void**)&B.dP, sizeof(Chunk) * std::max<size_t>(P.size(), 1), ctx->stream));
        HIP_TRY(dev_alloc((void**)&B.dC, sizeof(Chunk) * std::max<size_t>(Cc.size(), 1), ctx->stream));
    }
    HIP_TRY(hipMemcpyAsync(B.dT[0], t0.data(), sizeof(DevScene) * ns, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(B.dT[1], t1.data(), sizeof(DevScene) * ns, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(B.dP, P.data(), sizeof(Chunk) * P.size(), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(B.dC, Cc.data(), sizeof(Chunk) * Cc.size(), hipMemcpyHostToDevice, ctx->stream));
    if (B.nP != (int)P.size() || B.nC != (int)Cc.size() || B.maxCellChunks != maxCC) {
        // the captured segments hold the launch sizes
        for (auto& a : B.seg)
            for (auto& e : a)
                if (e) { (void)hipGraphExecDestroy(e); e = nullptr; }
    }
    B.nP = (int)P.size();
    B.nC = (int)Cc.size();
    B.maxCellChunks = maxCC;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SPHG_OK;
}

// Counting sort of the fluid into (cell, id) order; tables T (before) and
// T2 (after: the other half of each double buffer is current).  Keys and
// histogram are formed by the caller (k_keys or k_df_positions).
void grid_sort(sphg_ctx* ctx, Batch& B, const DevScene* T) {
    Profiler* P = &ctx->prof;
    hipStream_t st = ctx->stream;
    const int ns = (int)B.scenes.size();
    { Scope s(P, st, K_SCAN); launch_scan(st, T, B.dC, B.nC, ns, B.maxCellChunks); }
    { Scope s(P, st, K_SCATTER); launch_scatter(st, T, B.dP, B.nP); }
    { Scope s(P, st, K_REORDER); launch_rank_reorder(st, T, B.dP, B.nP); }
}

int read_active(sphg_ctx* ctx, int& v, int& maxIt, uint32_t& syncs) {
    HIP_TRY(hipMemcpyAsync(ctx->hActive, ctx->dActive, 2 * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    v = ctx->hActive[0];
    maxIt = ctx->hActive[1];
    ++syncs;
    return SPHG_OK;
}

void flip(Batch& B) {
    std::swap(B.dT[0], B.dT[1]);
    for (size_t i = 0; i < B.scenes.size(); ++i) {
        B.scenes[i]->par ^= 1;
        B.parity[i] = B.scenes[i]->par;
    }
}

// The one-scene tables of s, refreshed if another batch moved its parity.
int self_tables(sphg_scene* s) {
    Batch& B = *s->self;
    if (B.parity[0] != s->par || B.nAt[0] != s->n) return upload_batch_tables(s->ctx, B);
    return SPHG_OK;
}

// Finalize kernel of a reduction (one block per scene).
void finalize(sphg_ctx* ctx, Batch& B, const DevScene* T, int kind) {
    Scope s(&ctx->prof, ctx->stream, K_FINALIZE);
    launch_finalize(ctx->stream, kind, T, (int)B.scenes.size(), ctx->dActive);
}

// Whether iteration `it` of a loop is to be launched.  Below `ahead` --
// the batch's smallest minimum iteration count, or the iteration count the
// loop reached in the previous step -- it is launched without asking: a
// scene that has already converged skips it on the device.  Past that, the
// number of scenes still iterating (and the largest count reached) is read
// back.  `seen` returns the largest count reached once the loop is over.
int loop_more(sphg_ctx* ctx, int it, int ahead, bool& more, int& seen, uint32_t& syncs) {
    if (it < ahead) {
        more = true;
        return SPHG_OK;
    }
    int act = 0;
    int rc = read_active(ctx, act, seen, syncs);
    more = act > 0;
    return rc;
}

// Iterations of a density loop launched before the first read-back: the
// count the loop needed in the previous step plus one.  The count moves by
// about one between steps; an iteration that a converged scene does not need
// is three launches that return at once (~12 us), a read-back ~40-50 us of
// idle device.  The divergence loop (one iteration nearly always) gets no
// slack.
int density_ahead(const Batch& B) {
    return std::max(B.minIt, B.lastIt > 0 ? B.lastIt + B.slackD : 0);
}
int divergence_ahead(const Batch& B) {
    return std::max(B.minItV, B.lastItV > 0 ? B.lastItV + B.slackV : 0);
}

// A fixed launch sequence of the step (kernels only, no copies or
// memsets): launched as issued while kernel timing is on (its events need
// the single launches) or for a batch that lives for one call; otherwise
// captured once per table parity into a graph and replayed.
template <class F>
int run_segment(sphg_ctx* ctx, Batch& B, int which, F&& issue) {
    const bool replay = B.replay >= 0 ? B.replay != 0 : ctx->graphs;
    if (!B.graphs || !replay || ctx->prof.on) {
        issue();
        return SPHG_OK;
    }
    const int k = B.dT[0] == B.tab0 ? 0 : 1;
    hipGraphExec_t& e = B.seg[which][k];
    if (!e) {
        hipGraph_t g = nullptr;
        HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        issue();
        hipError_t ce = hipStreamEndCapture(ctx->stream, &g);
        if (ce == hipSuccess) ce = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
        if (ce != hipSuccess) {
            // nothing ran: launch the segment one kernel at a time, and do
            // not try to capture for this batch again
            e = nullptr;
            (void)hipGetLastError();
            B.graphs = false;
            issue();
            return SPHG_OK;
        }
    }
    HIP_TRY(hipGraphLaunch(e, ctx->stream));
    return SPHG_OK;
}

// One step of every scene in B (all of one method).
int batch_one_step(sphg_ctx* ctx, Batch& B, uint32_t& syncs) {
    Profiler* P = &ctx->prof;
    hipStream_t st = ctx->stream;
    const int method = B.scenes[0]->desc.method;
    const int ns = (int)B.scenes.size();
    int anyFlags = 0, anySt = 0, anyCube = 0, anyEmit = 0, anyPb = 0;
    for (auto* s : B.scenes) {
        anyFlags |= s->view[0].flags;
        anySt |= (s->view[0].flags & F_ST);
        anyCube |= s->view[0].cubeOn;
        anyEmit |= s->view[0].nEmit;
        anyPb |= s->view[0].pbOn;
    }
    // boundary-particle passes of the PCISPH boundary method, one launch per scene
    auto pbDensity = [&](const DevScene* Tt, int pred, int iter) {
        for (int k = 0; k < ns; ++k)
            if (B.scenes[k]->view[0].pbOn) launch_pb_bdensity(st, pred, Tt, k, (int)B.scenes[k]->nb, iter);
    };
    int rc;
    bool more = false;
    if (method == DFSPH) {
        const DevScene* T = B.dT[0];
        rc = run_segment(ctx, B, 0, [&] {
            if (anyFlags) {
                if (anySt) { Scope s(P, st, K_NORMALS); launch_normals(st, T, B.dP, B.nP); }
                Scope s(P, st, K_FORCES);
                launch_forces(st, DFSPH, T, B.dP, B.nP);
            }
            { Scope s(P, st, K_DT); launch_dt(st, T, B.dP, B.nP); }
            finalize(ctx, B, T, 0);
            { Scope s(P, st, K_DFPREDICT); launch_df_predict(st, T, B.dP, B.nP); }
            { Scope s(P, st, K_DFVEL_WARM); launch_df_vel(st, 0, T, B.dP, B.nP); }
            // the loop counters were cleared by the dt finalize
            { Scope s(P, st, K_DFERR_DENS); launch_df_err(st, 0, T, B.dP, B.nP); }
            finalize(ctx, B, T, 1);
        });
        if (rc) return rc;
        int seen = 0;
        for (int it = 0; it < B.maxIt; ++it) {
            if ((rc = loop_more(ctx, it, density_ahead(B), more, seen, syncs))) return rc;
            if (!more) break;
            { Scope s(P, st, K_DFVEL_DENS); launch_df_vel(st, 1, T, B.dP, B.nP); }
            { Scope s(P, st, K_DFERR_DENS); launch_df_err(st, 1, T, B.dP, B.nP); }
            finalize(ctx, B, T, 2);
        }
        B.lastIt = seen;
        const DevScene* T2 = B.dT[1];
        rc = run_segment(ctx, B, 1, [&] {
            { Scope s(P, st, K_DFPOS); launch_df_positions(st, T, B.dP, B.nP, ctx->dActive); }
            // PCISPH boundary method: projection after the position update
            // (src/DFSPHSolver.cpp:65-75), which then counts the cells
            if (anyPb) { Scope s(P, st, K_PB); launch_pb_correct(st, 0, T, B.dP, B.nP); }
            // emitters between the position update and the grid rebuild
            // (src/DFSPHSolver.cpp:77); the new particles' cells are counted
            if (anyEmit) { Scope s(P, st, K_EMIT); launch_emit(st, B.dT[0], B.dT[1], ns, 1); }
            grid_sort(ctx, B, T);
            // (the loop counters were cleared by the position update)
            // the build also forms the divergence solve's initial error,
            // kappa_v and the error partials (computeDivergenceError before
            // the loop)
            { Scope s(P, st, K_BUILD); launch_build(st, DFSPH, T2, B.dP, B.nP); }
            finalize(ctx, B, T2, 3);
        });
        if (rc) return rc;
        seen = 0;
        for (int it = 0; it < B.maxIt; ++it) {
            if ((rc = loop_more(ctx, it, divergence_ahead(B), more, seen, syncs))) return rc;
            if (!more) break;
            { Scope s(P, st, K_DFVEL_DIV); launch_df_vel(st, 2, T2, B.dP, B.nP); }
            { Scope s(P, st, K_DFERR_DIV); launch_df_err(st, 3, T2, B.dP, B.nP); }
            finalize(ctx, B, T2, 4);
        }
        B.lastItV = seen;
        // cube boundary after the divergence solve (src/DFSPHSolver.cpp:99-108)
        if (anyCube) { Scope s(P, st, K_CUBE); launch_cube(st, T2, B.dP, B.nP); }
        flip(B);
    } else if (method == WCSPH) {
        const DevScene* T = B.dT[0];
        const DevScene* T2 = B.dT[1];
        rc = run_segment(ctx, B, 0, [&] {
            // the keys kernel also starts the step (WCSPH has no loop to count)
            { Scope s(P, st, K_KEYS); launch_keys(st, T, B.dP, B.nP, ctx->dActive); }
            grid_sort(ctx, B, T);
            { Scope s(P, st, K_BUILD); launch_build(st, WCSPH, T2, B.dP, B.nP); }
            // PCISPH boundary method: boundary densities and pressures
            // (src/SPHSolver.cpp:123-135, src/WCSPHSolver.cpp:116-128)
            if (anyPb) { Scope s(P, st, K_PB); pbDensity(T2, 0, 0); }
            if (anySt) { Scope s(P, st, K_NORMALS); launch_normals(st, T2, B.dP, B.nP); }
            { Scope s(P, st, K_FORCES); launch_forces(st, WCSPH, T2, B.dP, B.nP); }
            // cube boundary after the integration (src/WCSPHSolver.cpp:51-60),
            // or the PCISPH boundary projection (:62-73), then the emitters
            // (:75), then time += dt (:77)
            if (anyCube) { Scope s(P, st, K_CUBE); launch_cube(st, T2, B.dP, B.nP); }
            if (anyPb) { Scope s(P, st, K_PB); launch_pb_correct(st, 1, T2, B.dP, B.nP); }
            if (anyEmit) { Scope s(P, st, K_EMIT); launch_emit(st, B.dT[1], B.dT[0], ns, 0); }
            // the forces advanced the time of scenes with none of these
            if (anyCube || anyPb || anyEmit) { Scope s(P, st, K_TIME); launch_time_advance(st, T2, ns); }
        });
        if (rc) return rc;
        flip(B);
    } else {
        const DevScene* T = B.dT[0];
        const DevScene* T2 = B.dT[1];
        rc = run_segment(ctx, B, 0, [&] {
            launch_fill_int(st, ctx->dActive, 0, 2);
            // the keys kernel also starts the step and counts the scenes
            // whose pressure loop runs
            { Scope s(P, st, K_KEYS); launch_keys(st, T, B.dP, B.nP, ctx->dActive); }
            grid_sort(ctx, B, T);
            { Scope s(P, st, K_BUILD); launch_build(st, PCISPH, T2, B.dP, B.nP); }
            if (anySt) { Scope s(P, st, K_NORMALS); launch_normals(st, T2, B.dP, B.nP); }
            { Scope s(P, st, K_FORCES); launch_forces(st, PCISPH, T2, B.dP, B.nP); }
        });
        if (rc) return rc;
        int seen = 0;
        for (int it = 0; it < B.maxIt; ++it) {
            if ((rc = loop_more(ctx, it, density_ahead(B), more, seen, syncs))) return rc;
            if (!more) break;
            // PCISPH boundary method: correctPredPositions after the
            // prediction (src/PCISPHSolver.cpp:370-377)
            if (anyPb) { Scope s(P, st, K_PB); launch_pb_predcorrect(st, T2, B.dP, B.nP); }
            { Scope s(P, st, K_PCIDENS); launch_pci_density(st, T2, B.dP, B.nP); }
            finalize(ctx, B, T2, 5);
            // boundary densities and pressures of the iteration (:151-162, :276-293)
            if (anyPb) { Scope s(P, st, K_PB); pbDensity(T2, 1, it); }
            { Scope s(P, st, K_PCIPACC); launch_pci_pacc(st, T2, B.dP, B.nP, it); }
        }
        B.lastIt = seen;
        rc = run_segment(ctx, B, 1, [&] {
            { Scope s(P, st, K_PCIINT); launch_pci_integrate(st, T2, B.dP, B.nP); }
            // cube boundary after the integration (src/PCISPHSolver.cpp:54-63),
            // or the PCISPH boundary projection (:65-76), then the emitters
            // (:78), then time += dt (:80)
            if (anyCube) { Scope s(P, st, K_CUBE); launch_cube(st, T2, B.dP, B.nP); }
            if (anyPb) { Scope s(P, st, K_PB); launch_pb_correct(st, 2, T2, B.dP, B.nP); }
            if (anyEmit) { Scope s(P, st, K_EMIT); launch_emit(st, B.dT[1], B.dT[0], ns, 0); }
            // the integration advanced the time of scenes with none of these
            if (anyCube || anyPb || anyEmit) { Scope s(P, st, K_TIME); launch_time_advance(st, T2, ns); }
        });
        if (rc) return rc;
        flip(B);
    }
    return SPHG_OK;
}

// `events`: time the steps with HIP events (false: with the host clock
// around the steps and a stream synchronisation, creating nothing).
int run_batch(sphg_ctx* ctx, Batch& B, uint32_t nsteps, sphg_step_stats* stats, bool events = true) {
    hipStream_t st = ctx->stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (events) {
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
    }
    std::vector<Scalars> before(B.scenes.size());
    if (stats) {
        for (size_t s = 0; s < B.scenes.size(); ++s)
            HIP_TRY(hipMemcpyAsync(&before[s], B.scenes[s]->ss, sizeof(Scalars), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    uint32_t syncs = 0;
    // the overflow maximum reports the lists this call's steps use
    launch_clear_overflow(st, B.dT[0], (int)B.scenes.size());
    const auto h0 = std::chrono::steady_clock::now();
    if (events) HIP_TRY(hipEventRecord(e0, st));
    for (uint32_t k = 0; k < nsteps; ++k) {
        int rc = batch_one_step(ctx, B, syncs);
        if (rc) {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            return rc;
        }
    }
    float ms = 0.f;
    if (events) {
        HIP_TRY(hipEventRecord(e1, st));
        HIP_TRY(hipEventSynchronize(e1));
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    } else {
        HIP_TRY(hipStreamSynchronize(st));
        ms = (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count();
    }
    unsigned overflow = 0, nover = 0;
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->gpu_ms = ms;
        stats->steps = nsteps;
        stats->host_syncs = syncs;
    }
    for (size_t s = 0; s < B.scenes.size(); ++s) {
        Scalars now;
        HIP_TRY(hipMemcpy(&now, B.scenes[s]->ss, sizeof(Scalars), hipMemcpyDeviceToHost));
        if (stats) {
            stats->iterations += now.sumIt - before[s].sumIt;
            stats->iterations_v += now.sumItV - before[s].sumItV;
        }
        overflow = std::max(overflow, now.overflow);
        nover += now.overflow ? 1u : 0u;
        sphg_scene* sc = B.scenes[s];
        sc->stepped = nsteps > 0 || sc->stepped;
        if (!sc->emitters.empty() && now.nActive != sc->n) {
            sc->n = now.nActive;
            sc->view[0].n = sc->view[1].n = (int)now.nActive;
            B.nAt[s] = now.nActive;
        }
    }
    if (overflow)
        return fail(SPHG_EOVERFLOW, std::to_string(nover) + " scene(s) had a particle with up to " +
                                        std::to_string(overflow) +
                                        " neighbors, above max_neighbors: their results of this call are invalid "
                                        "(sphg_scene_get_scalars names them: neighbor_overflow > 0); the other "
                                        "scenes of the batch stepped normally");
    return SPHG_OK;
}

// Launch-schedule autotuner.  Every candidate schedule gives the same
// results -- the slack only moves read-backs and launches that return at
// once for converged scenes, and replay launches the same kernels -- so the
// trials are ordinary steps of the simulation.  Each round runs every
// candidate for `per` steps in turn (so all of them see about the same
// physics), and the candidate with the least wall time summed over the
// rounds is kept; with two or more rounds the first is a warm-up (graph
// capture, caches) and is not counted.
int tune_batch(sphg_ctx* ctx, Batch& B, uint32_t per, uint32_t rounds, sphg_tune_result* out) {
    if (per == 0 || rounds == 0) return fail(SPHG_EINVAL, "steps_per_trial and rounds must be positive");
    const int method = B.scenes[0]->desc.method;
    struct Cand {
        int d, v, r;
        double ms;
    };
    const std::vector<int> ds = method == WCSPH ? std::vector<int>{B.slackD} : std::vector<int>{0, 1, 2, 3};
    const std::vector<int> vs = method == DFSPH ? std::vector<int>{0, 1} : std::vector<int>{B.slackV};
    const std::vector<int> rs = B.graphs ? std::vector<int>{0, 1} : std::vector<int>{0};
    std::vector<Cand> cs;
    for (int d : ds)
        for (int v : vs)
            for (int r : rs) cs.push_back(Cand{d, v, r, 0.0});
    for (uint32_t round = 0; round < rounds; ++round) {
        for (auto& c : cs) {
            B.slackD = c.d;
            B.slackV = c.v;
            B.replay = c.r;
            const auto t0 = std::chrono::steady_clock::now();
            const int rc = run_batch(ctx, B, per, nullptr);
            if (rc) return rc;
            const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            if (rounds < 2 || round > 0) c.ms += ms;
        }
    }
    const double counted = (double)per * (rounds < 2 ? rounds : rounds - 1);
    size_t best = 0;
    double def = -1.0;
    for (size_t k = 0; k < cs.size(); ++k) {
        if (cs[k].ms < cs[best].ms) best = k;
        if (cs[k].d == 1 && cs[k].v == 0 && cs[k].r == 0) def = cs[k].ms / counted;
    }
    B.slackD = cs[best].d;
    B.slackV = cs[best].v;
    B.replay = cs[best].r;
    if (out) {
        out->density_slack = B.slackD;
        out->divergence_slack = B.slackV;
        out->replay = B.replay;
        out->candidates = (uint32_t)cs.size();
        out->steps = per * rounds * (uint32_t)cs.size();
        out->ms_per_step = cs[best].ms / counted;
        out->default_ms_per_step = def;
    }
    return SPHG_OK;
}

// An emitter's group in local coordinates (src/Emitter.cpp:31-75), at
// particle spacing dist = 2r.  Circle: the centre, then for ring l = 1 ..
// floor((width - dist/2) / dist) of radius l dist, floor(2 pi l dist / dist)
// particles at angles i 2pi/count, placed (radius sin, radius cos, 0).
// Square: floor(width / dist) x floor(height / dist) particles, x index
// outer, at (W/2 - i dist, H/2 - j dist, 0) with W, H the lattice extents.
void emitter_group(const sphg_emitter_desc& e, double particleRadius, std::vector<double>& out) {
    const double dist = 2.0 * particleRadius;
    auto put = [&](double x, double y) { out.push_back(x); out.push_back(y); out.push_back(0.0); };
    if (e.type == SPHG_EMITTER_CIRCLE) {
        const double lv = std::floor((e.width - 0.5 * dist) / dist);
        const unsigned levels = lv > 0 ? (unsigned)lv : 0u;
        put(0.0, 0.0);
        for (unsigned l = 0; l < levels; ++l) {
            const double radius = (l + 1) * dist;
            const unsigned count = (unsigned)std::floor(2.0 * M_PI * radius / dist);
            const double inc = 2 * M_PI / count;
            for (unsigned i = 0; i < count; ++i) {
                const double angle = i * inc;
                put(radius * std::sin(angle), radius * std::cos(angle));
            }
        }
    } else {
        const unsigned nx = (unsigned)std::floor(e.width / dist);
        const unsigned ny = (unsigned)std::floor(e.height / dist);
        const double w = (double)(nx - 1) * dist, hgt = (double)(ny - 1) * dist;
        for (unsigned i = 0; i < nx; ++i)
            for (unsigned j = 0; j < ny; ++j) put(w * 0.5 - i * dist, hgt * 0.5 - j * dist);
    }
}

// Grid extent and device memory plan of a scene (one allocation).
struct Plan {
    int k0[3], dims[3];
    long ncells = 0;
    int nChunks = 0, nCellChunks = 0, maxNbr = 0, flags = 0;
    size_t npad = 0, total = 0;
    size_t oX[2], oV[2], oK[2] = {0, 0}, oKV[2] = {0, 0}, oID[2];
    size_t oVYZ, oD0, oA, oXK, oGB, oAUX, oXS, oPACC, oGBS, oNRM, oADH, oRHO, oALPHA, oDRHOS, oDIV, oP;
    size_t oCELLU, oCELL, oPERM, oCC, oCS, oCCS, oNBR, oNNB, oNSH, oNG, oGC, oXB, oBCS, oPART, oSS, oTMP;
    size_t oEMIT = 0, oGROUP = 0;
    size_t oBN = 0, oBSELF = 0, oBDENS = 0, oBPRES = 0, oBPR = 0, oPBN[2] = {0, 0}, oPV = 0, oPRS = 0, oPVS = 0,
           oX0S = 0, oOCC = 0;
    std::vector<double> group;          // every emitter's group, 3 doubles per particle
    std::vector<int> groupOff, groupSize;
};

int make_plan(const sphg_scene* s, Plan& P) {
    const sphg_scene_desc& d = s->desc;
    const double h = 4.0 * d.particle_radius;
    const int method = d.method;
    // grid bounds in key space
    double lo[3], hi[3];
    bool given = false;
    for (int a = 0; a < 3; ++a) given |= (d.domain_min[a] != 0.0 || d.domain_max[a] != 0.0);
    if (given) {
        for (int a = 0; a < 3; ++a) { lo[a] = d.domain_min[a]; hi[a] = d.domain_max[a]; }
    } else {
        for (int a = 0; a < 3; ++a) { lo[a] = 1e300; hi[a] = -1e300; }
        auto acc = [&](const std::vector<float>& p) {
            for (size_t i = 0; i + 2 < p.size(); i += 3)
                for (int a = 0; a < 3; ++a) {
                    const double v = p[i + a];
                    if (std::isfinite(v)) { lo[a] = std::min(lo[a], v); hi[a] = std::max(hi[a], v); }
                }
        };
        acc(s->fxyz);
        acc(s->bxyz);
    }
    int* k0 = P.k0;
    int* dims = P.dims;
    long ncells = 1;
    for (int a = 0; a < 3; ++a) {
        const long kl = (long)std::floor(lo[a] / h) - 1;
        const long kh = (long)std::floor(hi[a] / h) + 1;
        if (kh - kl + 3 > (1L << 20)) return fail(SPHG_EINVAL, "grid too large along one axis");
        k0[a] = (int)kl;
        dims[a] = (int)(kh - kl + 1) + 2;
        ncells *= dims[a];
    }
    if (ncells > (1L << 28)) return fail(SPHG_EINVAL, "grid has more than 2^28 cells; set domain_min/max");

    const uint32_t n = s->cap, nb = s->nb;
    const int nChunks = (int)((n + kBlock - 1) / kBlock);
    const size_t npad = (size_t)nChunks * kBlock;
    const int nCellChunks = (int)((ncells + 1 + kCellChunk - 1) / kCellChunk);
    // a multiple of 4: the lists are padded to multiples of 4 per group
    const int maxNbr = ((d.max_neighbors ? (int)d.max_neighbors : kDefaultMaxNbr) + 3) & ~3;
    // the kernels address particle arrays with 32-bit byte offsets and the
    // neighbor lists with 32-bit element offsets
    if (npad > (size_t(1) << 28))
        return fail(SPHG_EINVAL, "more than 2^28 particle slots in one scene");
    if ((uint64_t)npad * (uint64_t)maxNbr > 0xFFFFFFFFull)
        return fail(SPHG_EINVAL, "particle slots x max_neighbors above 2^32 - 1 (" + std::to_string(npad) + " x " +
                                     std::to_string(maxNbr) + "): lower max_neighbors or split the scene");

    int flags = 0;
    // a viscosity force exists iff its coefficient is nonzero
    // (src/Simulation.cpp:423-424)
    if (d.viscosity != 0.0) {
        if (d.viscosity_method == SPHG_VISCOSITY_XSPH) flags |= F_XSPH;
        if (d.viscosity_method == SPHG_VISCOSITY_STANDARD) flags |= F_VSTD;
        if (d.viscosity_method == SPHG_VISCOSITY_ARTIFICIAL) flags |= F_VART;
    }
    if (d.surface_tension_method == SPHG_SURFACE_TENSION_AKINCI && d.surface_tension != 0.0) flags |= F_ST;
    if (d.adhesion_method == SPHG_ADHESION_AKINCI && d.adhesion != 0.0) flags |= F_ADH;

    // memory plan
    Arena A;
    const size_t f4 = sizeof(float4) * npad, f1 = sizeof(float) * npad, u1 = sizeof(uint32_t) * npad;
    size_t* oX = P.oX;
    size_t* oV = P.oV;
    size_t* oK = P.oK;
    size_t* oKV = P.oKV;
    size_t* oID = P.oID;
    for (int p = 0; p < 2; ++p) { oX[p] = A.reserve(f4); oV[p] = A.reserve(f4); oID[p] = A.reserve(u1); }
    if (method == DFSPH)
        for (int p = 0; p < 2; ++p) { oK[p] = A.reserve(f1); oKV[p] = A.reserve(f1); }
    P.oA = A.reserve(f4);
    P.oXK = A.reserve(f4);
    P.oVYZ = (method == DFSPH) ? A.reserve(2 * f1) : 0;
    P.oD0 = (method == DFSPH) ? A.reserve(sizeof(double) * npad) : 0;
    P.oGB = A.reserve(f4);
    P.oAUX = (method != DFSPH) ? A.reserve(f4) : 0;
    P.oXS = (method == PCISPH) ? A.reserve(f4) : 0;
    P.oPACC = (method == PCISPH) ? A.reserve(f4) : 0;
    P.oGBS = (method == PCISPH) ? A.reserve(f4) : 0;
    P.oNRM = (flags & F_ST) ? A.reserve(f4) : 0;
    P.oADH = (flags & F_ADH) ? A.reserve(f4) : 0;
    P.oRHO = A.reserve(f1);
    P.oALPHA = (method == DFSPH) ? A.reserve(f1) : 0;
    P.oDRHOS = (method == DFSPH) ? A.reserve(f1) : 0;
    P.oDIV = (method == DFSPH) ? A.reserve(f1) : 0;
    P.oP = (method != DFSPH) ? A.reserve(f1) : 0;
    P.oCELLU = A.reserve(u1);
    P.oCELL = A.reserve(u1);
    P.oPERM = A.reserve(u1);
    P.oCC = A.reserve(sizeof(int) * (size_t)ncells);
    P.oCS = A.reserve(sizeof(int) * (size_t)(ncells + 1));
    P.oCCS = A.reserve(sizeof(int) * (size_t)nCellChunks);
    P.oNBR = A.reserve(sizeof(uint32_t) * npad * (size_t)maxNbr);
    P.oNNB = A.reserve(sizeof(uint16_t) * npad);
    P.oNSH = (method == PCISPH) ? A.reserve(sizeof(uint16_t) * npad) : 0;
    P.oNG = A.reserve(npad);
    P.oGC = A.reserve(sizeof(uint16_t) * (npad / kGroup));
    P.oXB = A.reserve(sizeof(float4) * std::max<size_t>(nb, 1));
    P.oBCS = A.reserve(sizeof(int) * (size_t)(ncells + 1));
    P.oPART = A.reserve(sizeof(double) * (size_t)std::max(nChunks, 1));
    P.oSS = A.reserve(sizeof(Scalars));
    P.oTMP = A.reserve(std::max(f4, sizeof(float4) * std::max<size_t>(nb, 1)));
    if (d.boundary_method == SPHG_BOUNDARY_PCISPH) {
        const size_t nb1 = std::max<size_t>(nb, 1);
        P.oBN = A.reserve(sizeof(float4) * nb1);
        P.oBSELF = A.reserve(sizeof(double) * 2 * nb1);
        P.oBDENS = A.reserve(sizeof(float) * nb1);
        P.oBPRES = A.reserve(sizeof(float) * nb1);
        P.oBPR = A.reserve(sizeof(float) * nb1);
        P.oPBN[0] = A.reserve(f4);
        P.oPBN[1] = (method == DFSPH) ? A.reserve(f4) : 0;
        if (method == PCISPH) {
            P.oPV = A.reserve(f4);
            P.oPRS = A.reserve(f4);
            P.oPVS = A.reserve(f4);
            P.oX0S = A.reserve(f4);
        }
        P.oOCC = A.reserve((size_t)ncells);
    }
    if (!s->emitters.empty()) {
        P.group.clear();
        for (const auto& e : s->emitters) {
            P.groupOff.push_back((int)(P.group.size() / 3));
            const size_t before = P.group.size();
            emitter_group(e, d.particle_radius, P.group);
            P.groupSize.push_back((int)((P.group.size() - before) / 3));
        }
        P.oEMIT = A.reserve(sizeof(EmitterDev) * s->emitters.size());
        P.oGROUP = A.reserve(sizeof(double) * std::max<size_t>(P.group.size(), 1));
    }
    P.ncells = ncells;
    P.nChunks = nChunks;
    P.nCellChunks = nCellChunks;
    P.maxNbr = maxNbr;
    P.flags = flags;
    P.npad = npad;
    P.total = A.used;
    return SPHG_OK;
}

int check_batch(sphg_ctx* ctx, sphg_scene* const* scenes, uint32_t nscenes) {
    const int method = scenes[0] ? scenes[0]->desc.method : -1;
    for (uint32_t i = 0; i < nscenes; ++i) {
        if (!scenes[i] || scenes[i]->ctx != ctx) return fail(SPHG_EINVAL, "scene of another context");
        if (!scenes[i]->ready) return fail(SPHG_ESTATE, "scene not initialised");
        if (scenes[i]->desc.method != method) return fail(SPHG_EINVAL, "a batch steps scenes of one method");
        for (uint32_t j = 0; j < i; ++j)
            if (scenes[j] == scenes[i]) return fail(SPHG_EINVAL, "scene listed twice");
    }
    return SPHG_OK;
}

// This device's address of `scratch`, looked up once per device (with the
// device current) and kept.
std::mutex g_scratch_mu;
std::map<int, unsigned char*> g_scratch_addr;

int scratch_address(int device, unsigned char** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch_addr.find(device);
    if (it != g_scratch_addr.end()) {
        *out = it->second;
        return SPHG_OK;
    }
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(SPHG_EHIP, std::string("address of the allocation-free step's scratch array: ") +
                                   hipGetErrorString(e));
    g_scratch_addr[device] = (unsigned char*)p;
    *out = (unsigned char*)p;
    return SPHG_OK;
}

}  // namespace

// ================================================================ C API

struct sphg_batch {
    sphg_ctx* ctx = nullptr;
    Batch B;
};

extern "C" {

const char* sphg_version(void) { return "sphg 0.1 (gfx950)"; }
const char* sphg_last_error(void) { return g_err.c_str(); }

int sphg_ctx_create(int32_t device, sphg_ctx** out) {
    if (!out) return fail(SPHG_EINVAL, "out is NULL");
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(SPHG_EINVAL, "no HIP device " + std::to_string(device));
    HIP_TRY(hipSetDevice(device));
    sphg_ctx* c = new sphg_ctx();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) {
        // freed scene memory stays in the device's default pool for the next
        // scene instead of being released at the next synchronisation
        hipMemPool_t pool = nullptr;
        if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess && pool) {
            uint64_t keep = UINT64_MAX;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        }
    }
    if (e == hipSuccess) e = hipMalloc(&c->dActive, 2 * sizeof(int));
    if (e == hipSuccess) e = hipHostMalloc(&c->hActive, 2 * sizeof(int), hipHostMallocDefault);
    if (e != hipSuccess) {
        const std::string msg = std::string("context set-up: ") + hipGetErrorString(e);
        (void)sphg_ctx_destroy(c);   // releases the streams and buffers made so far
        return fail(SPHG_EHIP, msg);
    }
    const int rc = scratch_address(device, &c->scratch);
    if (rc) {
        (void)sphg_ctx_destroy(c);
        return rc;
    }
    *out = c;
    return SPHG_OK;
}

int sphg_ctx_destroy(sphg_ctx* ctx) {
    if (!ctx) return SPHG_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->copyStream) (void)hipStreamSynchronize(ctx->copyStream);
    {
        std::lock_guard<std::mutex> lk(ctx->scenesMu);
        for (auto* s : ctx->scenes) {
            s->ctx = nullptr;
            delete s;
        }
        ctx->scenes.clear();
    }
    ctx->prof.collect();
    if (ctx->dActive) (void)hipFree(ctx->dActive);
    if (ctx->hActive) (void)hipHostFree(ctx->hActive);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    if (ctx->copyStream) (void)hipStreamDestroy(ctx->copyStream);
    delete ctx;
    return SPHG_OK;
}

int sphg_ctx_graphs(sphg_ctx* ctx, int32_t enable) {
    if (!ctx) return fail(SPHG_EINVAL, "ctx is NULL");
    ctx->graphs = enable != 0;
    return SPHG_OK;
}

int sphg_ctx_synchronize(sphg_ctx* ctx) {
    if (!ctx) return fail(SPHG_EINVAL, "ctx is NULL");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SPHG_OK;
}

int sphg_scene_create(sphg_ctx* ctx, const sphg_scene_desc* desc, sphg_scene** out) {
    if (!ctx || !desc || !out) return fail(SPHG_EINVAL, "NULL argument");
    int rc = check_desc(*desc);
    if (rc) return rc;
    sphg_scene* s = new sphg_scene();
    s->ctx = ctx;
    s->desc = *desc;
    {
        std::lock_guard<std::mutex> lk(ctx->scenesMu);
        ctx->scenes.push_back(s);
    }
    *out = s;
    return SPHG_OK;
}

int sphg_scene_destroy(sphg_scene* s) {
    if (!s) return SPHG_OK;
    if (s->ctx) {
        // a scene that was never given device memory has no work to wait for
        if (s->ready || s->mem) {
            (void)hipSetDevice(s->ctx->device);
            (void)hipStreamSynchronize(s->ctx->stream);
        }
        std::lock_guard<std::mutex> lk(s->ctx->scenesMu);
        auto& v = s->ctx->scenes;
        v.erase(std::remove(v.begin(), v.end(), s), v.end());
    }
    delete s;
    return SPHG_OK;
}

int sphg_scene_set_fluid(sphg_scene* s, const float* xyz, const float* vel, uint32_t n) {
    if (!s || (!xyz && n)) return fail(SPHG_EINVAL, "NULL argument");
    if (s->ready) return fail(SPHG_ESTATE, "set_fluid after init: use sphg_scene_set_state");
    if (n == 0) return fail(SPHG_EINVAL, "a scene needs at least one fluid particle");
    s->fxyz.assign(xyz, xyz + 3 * (size_t)n);
    if (vel) s->fvel.assign(vel, vel + 3 * (size_t)n);
    else s->fvel.assign(3 * (size_t)n, 0.f);
    uint64_t cap = n;
    for (const auto& e : s->emitters) cap += e.num_particles;
    if (cap > (uint64_t)INT32_MAX) return fail(SPHG_EINVAL, "more than 2^31 - 1 particle slots");
    s->n = n;
    s->cap = (uint32_t)cap;
    return SPHG_OK;
}

int sphg_scene_add_emitter(sphg_scene* s, const sphg_emitter_desc* e) {
    if (!s || !e) return fail(SPHG_EINVAL, "NULL argument");
    if (s->ready) return fail(SPHG_ESTATE, "add_emitter after init");
    if (e->type != SPHG_EMITTER_SQUARE && e->type != SPHG_EMITTER_CIRCLE)
        return fail(SPHG_EINVAL, "emitter type must be 0 (square) or 1 (circle)");
    bool finite = std::isfinite(e->velocity) && std::isfinite(e->start_time) && std::isfinite(e->width) &&
                  std::isfinite(e->height) && std::isfinite(e->spacing);
    for (int c = 0; c < 3; ++c) finite &= std::isfinite(e->position[c]);
    for (int c = 0; c < 4; ++c) finite &= std::isfinite(e->rotation[c]);
    if (!finite) return fail(SPHG_EINVAL, "emitter values must be finite");
    const double dist = 2.0 * s->desc.particle_radius;
    if (e->width / dist > 1e6 || e->height / dist > 1e6) return fail(SPHG_EINVAL, "emitter group too large");
    const uint64_t cap = (uint64_t)s->cap + e->num_particles;
    if (cap > (uint64_t)INT32_MAX) return fail(SPHG_EINVAL, "more than 2^31 - 1 particle slots");
    s->emitters.push_back(*e);
    s->cap = (uint32_t)cap;
    return SPHG_OK;
}

int sphg_scene_capacity(sphg_scene* s, uint32_t* capacity) {
    if (!s || !capacity) return fail(SPHG_EINVAL, "NULL argument");
    *capacity = s->cap;
    return SPHG_OK;
}

int sphg_scene_set_boundary(sphg_scene* s, const float* xyz, uint32_t nb) {
    if (!s || (!xyz && nb)) return fail(SPHG_EINVAL, "NULL argument");
    if (s->ready) return fail(SPHG_ESTATE, "set_boundary after init");
    if (nb && s->desc.boundary_method != SPHG_BOUNDARY_AKINCI && s->desc.boundary_method != SPHG_BOUNDARY_PCISPH)
        return fail(SPHG_EINVAL, "boundary particles need boundaryMethod 1 (PCISPH) or 2 (Akinci)");
    // boundary indices share the neighbor lists and the int cell tables
    if (nb > (1u << 28)) return fail(SPHG_EINVAL, "more than 2^28 boundary particles in one scene");
    s->bxyz.assign(xyz, xyz + 3 * (size_t)nb);
    s->nb = nb;
    s->bnrm.clear();
    return SPHG_OK;
}

int sphg_scene_set_boundary_normals(sphg_scene* s, const float* nrm, uint32_t nb) {
    if (!s || (!nrm && nb)) return fail(SPHG_EINVAL, "NULL argument");
    if (s->ready) return fail(SPHG_ESTATE, "set_boundary_normals after init");
    if (s->desc.boundary_method != SPHG_BOUNDARY_PCISPH)
        return fail(SPHG_EINVAL, "boundary normals belong to boundaryMethod 1 (PCISPH)");
    if (nb != s->nb) return fail(SPHG_EINVAL, "one normal per boundary particle");
    for (size_t i = 0; i < 3 * (size_t)nb; ++i)
        if (!std::isfinite(nrm[i])) return fail(SPHG_EINVAL, "boundary normals must be finite");
    s->bnrm.assign(nrm, nrm + 3 * (size_t)nb);
    return SPHG_OK;
}

int sphg_scene_set_pb_normals(sphg_scene* s, const float* nrm) {
    if (!s || !nrm) return fail(SPHG_EINVAL, "NULL argument");
    if (!s->ready) return fail(SPHG_ESTATE, "set_pb_normals before init");
    if (!s->view[0].pbOn) return fail(SPHG_EINVAL, "fluid normals belong to boundaryMethod 1 (PCISPH)");
    sphg_ctx* ctx = s->ctx;
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t n = s->n;
    if (!n) return SPHG_OK;
    DevTemp d(ctx->stream);
    HIP_TRY(d.alloc(sizeof(float) * 3 * n));
    HIP_TRY(hipMemcpyAsync(d.p, nrm, sizeof(float) * 3 * n, hipMemcpyHostToDevice, ctx->stream));
    const DevScene& V = s->view[s->par];
    launch_gather_f3(ctx->stream, V.pbn, d.f(), V.id, (int)n);
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    HIP_TRY(hipGetLastError());
    return SPHG_OK;
}

int sphg_scene_init(sphg_scene* s) {
    if (!s || !s->ctx) return fail(SPHG_EINVAL, "NULL scene");
    if (s->ready) return fail(SPHG_ESTATE, "scene already initialised");
    if (s->cap == 0) return fail(SPHG_ESTATE, "no fluid particles and no emitters");
    if (s->desc.boundary_method == SPHG_BOUNDARY_PCISPH && s->bnrm.size() != 3 * (size_t)s->nb)
        return fail(SPHG_ESTATE, "the PCISPH boundary method needs sphg_scene_set_boundary_normals");
    sphg_ctx* ctx = s->ctx;
    HIP_TRY(hipSetDevice(ctx->device));
    const sphg_scene_desc& d = s->desc;
    const double r = d.particle_radius;
    const double h = 4.0 * r;
    const double rho0 = d.density0;
    const double m = rho0 * std::pow(2.0 * r, 3.0);
    const int method = d.method;

    Plan P;
    {
        int prc = make_plan(s, P);
        if (prc) return prc;
    }
    const int* k0 = P.k0;
    const int* dims = P.dims;
    const long ncells = P.ncells;
    const uint32_t n = s->n, nb = s->nb;
    const int nChunks = P.nChunks, nCellChunks = P.nCellChunks, maxNbr = P.maxNbr, flags = P.flags;
    hipError_t he = dev_alloc((void**)&s->mem, P.total, ctx->stream);
    s->memStream = ctx->stream;
    if (he != hipSuccess) return fail(SPHG_ENOMEM, "device allocation of " + std::to_string(P.total) + " bytes failed");
    s->memBytes = P.total;
    char* b = s->mem;
    HIP_TRY(hipMemsetAsync(b, 0, P.total, ctx->stream));

    DevScene D{};
    fill_coef(D.kc, h);
    fill_coef_d(D.kd, h);
    D.hD = h;
    D.r = r;
    D.cfl = d.cfl_factor;
    const double minFps = 1.0 / d.fps;
    s->dtMin = (minFps < d.min_time_step) ? minFps : d.min_time_step;   // src/TimeManager.h:55
    s->dtMax = d.max_time_step;
    D.dtMin = s->dtMin;
    D.dtMax = s->dtMax;
    D.eta = d.eta;
    D.etaV = d.eta_v;
    D.m = (float)m;
    D.rho0 = (float)rho0;
    D.mD = m;
    D.rho0D = rho0;
    D.W0 = D.kc.ck;
    D.pciDeltaD = (method == PCISPH) ? pcisph_delta(r, h, m, rho0, d.time_step) : 0.0;
    D.pciDelta = (float)D.pciDeltaD;
    D.gx = (float)d.gravity[0];
    D.gy = (float)d.gravity[1];
    D.gz = (float)d.gravity[2];
    D.nu = (float)d.viscosity;
    D.nub = (float)d.boundary_viscosity;
    D.stCoef = (float)d.surface_tension;
    D.adhBeta = (float)d.adhesion;
    D.stiffness = (float)d.stiffness;
    D.gammaEos = (float)d.gamma;
    D.method = method;
    D.flags = flags;
    D.cubeOn = d.cube_enabled ? 1 : 0;
    for (int c = 0; c < 3; ++c) {
        float lo = (float)d.cube_min[c], hi = (float)d.cube_max[c];
        if ((double)lo < d.cube_min[c]) lo = std::nextafter(lo, INFINITY);
        if ((double)hi > d.cube_max[c]) hi = std::nextafter(hi, -INFINITY);
        D.cubeLo[c] = lo;
        D.cubeHi[c] = hi;
    }
    D.cubeNf = d.cube_normal_fct;
    D.cubeTf = d.cube_tangential_fct;
    D.pbOn = d.boundary_method == SPHG_BOUNDARY_PCISPH ? 1 : 0;
    D.pbNf = d.pb_normal_fct;
    D.pbTf = d.pb_tangential_fct;
    // PCISPH predicted densities from the list and a shell list instead of
    // the 27-cell walk (not with the PCISPH boundary method, whose collision
    // projection also moves the predicted positions)
    if (method == PCISPH && !D.pbOn) D.pciShell = 1;
    // the viscosity forces' boundary terms exist for Akinci boundaries only
    // (src/XSPHViscosity.h:58, src/StandardViscosityForce.h:62, src/ArtificialViscosity.h:66)
    if (D.pbOn) D.nub = 0.f;
    D.minIt = (int)d.min_iterations;
    D.maxIt = (int)d.max_iterations;
    D.minItV = (int)d.min_iterations_v;
    D.maxItV = (int)d.max_iterations_v;
    D.kx0 = k0[0]; D.ky0 = k0[1]; D.kz0 = k0[2];
    D.nx = dims[0]; D.ny = dims[1]; D.nz = dims[2];
    D.ncells = (int)ncells;
    D.n = (int)n;
    D.cap = (int)s->cap;
    D.nb = (int)nb;
    D.maxNbr = maxNbr;
    D.nChunks = nChunks;
    D.nCellChunks = nCellChunks;
    auto F4 = [&](size_t o) { return (float4*)(b + o); };
    auto F1 = [&](size_t o) { return (float*)(b + o); };
    D.a = F4(P.oA);
    D.xk = F4(P.oXK);
    D.vyz = P.oVYZ ? (float2*)(b + P.oVYZ) : nullptr;
    D.d0 = P.oD0 ? (double*)(b + P.oD0) : nullptr;
    D.gb = F4(P.oGB);
    D.aux = P.oAUX ? F4(P.oAUX) : nullptr;
    D.xs = P.oXS ? F4(P.oXS) : nullptr;
    D.pacc = P.oPACC ? F4(P.oPACC) : nullptr;
    D.gbs = P.oGBS ? F4(P.oGBS) : nullptr;
    D.nrm = P.oNRM ? F4(P.oNRM) : nullptr;
    D.adh = P.oADH ? F4(P.oADH) : nullptr;
    D.drho = F1(P.oRHO);
    D.alpha = P.oALPHA ? F1(P.oALPHA) : nullptr;
    D.drhoS = P.oDRHOS ? F1(P.oDRHOS) : nullptr;
    D.divErr = P.oDIV ? F1(P.oDIV) : nullptr;
    D.p = P.oP ? F1(P.oP) : nullptr;
    D.cellU = (int*)(b + P.oCELLU);
    D.cell = (int*)(b + P.oCELL);
    D.perm = (uint32_t*)(b + P.oPERM);
    D.cellCount = (int*)(b + P.oCC);
    D.cellStart = (int*)(b + P.oCS);
    D.cellChunkSum = (int*)(b + P.oCCS);
    D.nbr = (uint32_t*)(b + P.oNBR);
    D.nnb = (uint16_t*)(b + P.oNNB);
    D.nsh = P.oNSH ? (uint16_t*)(b + P.oNSH) : nullptr;
    D.ngrad = (uint8_t*)(b + P.oNG);
    D.gcnt = (uint16_t*)(b + P.oGC);
    D.xb = F4(P.oXB);
    D.bCellStart = (int*)(b + P.oBCS);
    D.partial = (double*)(b + P.oPART);
    D.ss = (Scalars*)(b + P.oSS);
    s->ss = D.ss;
    if (D.pbOn) {
        D.bn = F4(P.oBN);
        D.bself = (double*)(b + P.oBSELF);
        D.bdens = F1(P.oBDENS);
        D.bpres = F1(P.oBPRES);
        D.bpr = F1(P.oBPR);
        D.pv = P.oPV ? F4(P.oPV) : nullptr;
        D.prS = P.oPRS ? F4(P.oPRS) : nullptr;
        D.pvS = P.oPVS ? F4(P.oPVS) : nullptr;
        D.x0s = P.oX0S ? F4(P.oX0S) : nullptr;
        D.everOcc = (uint8_t*)(b + P.oOCC);
    }
    if (!s->emitters.empty()) {
        // emitter constants and state, group table (src/Emitter.cpp:10-29)
        std::vector<EmitterDev> ed(s->emitters.size());
        for (size_t e = 0; e < s->emitters.size(); ++e) {
            const sphg_emitter_desc& E = s->emitters[e];
            EmitterDev& o = ed[e];
            for (int c = 0; c < 3; ++c) o.pos[c] = E.position[c];
            for (int c = 0; c < 4; ++c) o.q[c] = E.rotation[c];
            o.vel = E.velocity;
            o.interval = E.spacing * 1.105 * (2.0 * r) / E.velocity;   // src/Emitter.cpp:131,159
            o.next = E.start_time;
            o.remaining = E.num_particles;
            o.group = P.groupOff[e];
            o.groupSize = P.groupSize[e];
        }
        D.nEmit = (int)ed.size();
        D.emit = (EmitterDev*)(b + P.oEMIT);
        D.emitGroup = (const double*)(b + P.oGROUP);
        HIP_TRY(hipMemcpyAsync(D.emit, ed.data(), sizeof(EmitterDev) * ed.size(), hipMemcpyHostToDevice, ctx->stream));
        if (!P.group.empty())
            HIP_TRY(hipMemcpyAsync((void*)D.emitGroup, P.group.data(), sizeof(double) * P.group.size(),
                                   hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    s->tmp = F1(P.oTMP);
    for (int p = 0; p < 2; ++p) {
        DevScene v = D;
        v.x = F4(P.oX[p]); v.x2 = F4(P.oX[1 - p]);
        v.v = F4(P.oV[p]); v.v2 = F4(P.oV[1 - p]);
        v.id = (uint32_t*)(b + P.oID[p]); v.id2 = (uint32_t*)(b + P.oID[1 - p]);
        v.k = P.oK[p] ? F1(P.oK[p]) : nullptr; v.k2 = P.oK[1 - p] ? F1(P.oK[1 - p]) : nullptr;
        v.kv = P.oKV[p] ? F1(P.oKV[p]) : nullptr; v.kv2 = P.oKV[1 - p] ? F1(P.oKV[1 - p]) : nullptr;
        if (D.pbOn) {
            const size_t a0 = P.oPBN[1] ? P.oPBN[p] : P.oPBN[0], a1 = P.oPBN[1] ? P.oPBN[1 - p] : P.oPBN[0];
            v.pbn = F4(a0);
            v.pbn2 = F4(a1);
        }
        s->view[p] = v;
    }
    s->par = 0;

    // boundary: stable sort by linear cell on the host, cell table, upload
    {
        std::vector<std::pair<long, uint32_t>> keyed(nb);
        for (uint32_t i = 0; i < nb; ++i) {
            long lin = 0, mul = 1;
            for (int a = 0; a < 3; ++a) {
                double f = std::floor((double)s->bxyz[3 * i + a] / h);
                if (!(f == f)) f = k0[a];
                long ix = (long)std::min(std::max(f - k0[a] + 1.0, 1.0), (double)(dims[a] - 2));
                lin += ix * mul;
                mul *= dims[a];
            }
            keyed[i] = {lin, i};
        }
        std::stable_sort(keyed.begin(), keyed.end(), [](auto& x, auto& y) { return x.first < y.first; });
        std::vector<float4> xb(std::max<size_t>(nb, 1)), bn;
        std::vector<int> bcs((size_t)ncells + 1, 0);
        std::vector<uint8_t> occ;
        // PCISPH boundary method: V_b = (2r)^3, so rho0 V_b is the boundary
        // mass rho0 (2r)^3 (src/Simulation.cpp:465-467); its particles' cells
        // exist in the hash table from the start (src/SPHSolver.cpp:26)
        const float vPb = (float)std::pow(2.0 * r, 3.0);
        if (D.pbOn) {
            bn.resize(std::max<size_t>(nb, 1));
            occ.assign((size_t)ncells, 0);
        }
        s->border.resize(nb);
        for (uint32_t i = 0; i < nb; ++i) {
            const uint32_t o = keyed[i].second;
            s->border[i] = o;
            xb[i] = make_float4(s->bxyz[3 * o], s->bxyz[3 * o + 1], s->bxyz[3 * o + 2], D.pbOn ? vPb : 0.f);
            bcs[keyed[i].first + 1] += 1;
            if (D.pbOn) {
                bn[i] = make_float4(s->bnrm[3 * o], s->bnrm[3 * o + 1], s->bnrm[3 * o + 2], 0.f);
                occ[(size_t)keyed[i].first] = 1;
            }
        }
        for (long c = 0; c < ncells; ++c) bcs[c + 1] += bcs[c];
        HIP_TRY(hipMemcpyAsync(D.xb, xb.data(), sizeof(float4) * nb, hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(hipMemcpyAsync(D.bCellStart, bcs.data(), sizeof(int) * bcs.size(), hipMemcpyHostToDevice, ctx->stream));
        if (D.pbOn) {
            HIP_TRY(hipMemcpyAsync(D.bn, bn.data(), sizeof(float4) * nb, hipMemcpyHostToDevice, ctx->stream));
            HIP_TRY(hipMemcpyAsync(D.everOcc, occ.data(), occ.size(), hipMemcpyHostToDevice, ctx->stream));
        }
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }

    // scalars
    Scalars ss0{};
    ss0.dt = d.time_step;
    ss0.nActive = n;
    HIP_TRY(hipMemcpyAsync(D.ss, &ss0, sizeof(Scalars), hipMemcpyHostToDevice, ctx->stream));

    // tables of a one-scene batch
    s->self = new Batch();
    s->self->graphs = true;
    s->self->scenes = {s};
    int rc = upload_batch_tables(ctx, *s->self);
    if (rc) return rc;

    // boundary volumes (src/SPHSolver.cpp:22-39); the PCISPH method's
    // boundary-to-boundary density sums instead
    if (nb && !D.pbOn) launch_bvolume(ctx->stream, s->self->dT[0], 0, (int)nb, s->tmp);
    if (nb && D.pbOn) launch_pb_bself(ctx->stream, s->self->dT[0], 0, (int)nb);

    // fluid upload (staging in tmp is too small for xyz+vel: copy through a scratch allocation)
    if (n) {
        DevTemp stage(ctx->stream);
        HIP_TRY(stage.alloc(sizeof(float) * 6 * (size_t)n));
        float* dstage = stage.f();
        HIP_TRY(hipMemcpyAsync(dstage, s->fxyz.data(), sizeof(float) * 3 * n, hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(hipMemcpyAsync(dstage + 3 * (size_t)n, s->fvel.data(), sizeof(float) * 3 * n, hipMemcpyHostToDevice,
                               ctx->stream));
        launch_load_state(ctx->stream, s->self->dT[0], 0, (int)n, dstage, dstage + 3 * (size_t)n, nullptr, nullptr);
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }

    if (method == DFSPH) {
        // DFSPHSolver::init: insert, neighborhoodSearch, densities, alpha
        Batch& B = *s->self;
        launch_keys(ctx->stream, B.dT[0], B.dP, B.nP);
        grid_sort(ctx, B, B.dT[0]);
        launch_build(ctx->stream, DFSPH, B.dT[1], B.dP, B.nP);
        flip(B);
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    HIP_TRY(hipGetLastError());
    s->ready = true;
    s->fxyz.clear();
    s->fxyz.shrink_to_fit();
    s->fvel.clear();
    s->fvel.shrink_to_fit();
    return SPHG_OK;
}

int sphg_scene_set_state(sphg_scene* s, const float* xyz, const float* vel, const float* k, const float* kv,
                         double time_step, uint32_t last_iterations) {
    if (!s || !xyz || !vel) return fail(SPHG_EINVAL, "NULL argument");
    if (!s->ready) return fail(SPHG_ESTATE, "set_state before init");
    sphg_ctx* ctx = s->ctx;
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t n = s->n;
    DevTemp stage(ctx->stream);
    HIP_TRY(stage.alloc(sizeof(float) * 8 * n));
    float* dstage = stage.f();
    HIP_TRY(hipMemcpyAsync(dstage, xyz, sizeof(float) * 3 * n, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(dstage + 3 * n, vel, sizeof(float) * 3 * n, hipMemcpyHostToDevice, ctx->stream));
    if (k) HIP_TRY(hipMemcpyAsync(dstage + 6 * n, k, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
    if (kv) HIP_TRY(hipMemcpyAsync(dstage + 7 * n, kv, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
    Batch& B = *s->self;
    int rc = self_tables(s);
    if (rc) return rc;
    launch_load_state(ctx->stream, B.dT[0], 0, (int)n, dstage, dstage + 3 * n, k ? dstage + 6 * n : nullptr,
                      kv ? dstage + 7 * n : nullptr);
    Scalars ss;
    HIP_TRY(hipMemcpyAsync(&ss, s->ss, sizeof(Scalars), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ss.dt = time_step;
    ss.it = (int)last_iterations;
    s->stepped = false;
    HIP_TRY(hipMemcpyAsync(s->ss, &ss, sizeof(Scalars), hipMemcpyHostToDevice, ctx->stream));
    if (s->desc.method == DFSPH) {
        launch_keys(ctx->stream, B.dT[0], B.dP, B.nP);
        grid_sort(ctx, B, B.dT[0]);
        launch_build(ctx->stream, DFSPH, B.dT[1], B.dP, B.nP);
        flip(B);
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    HIP_TRY(hipGetLastError());
    return SPHG_OK;
}

int sphg_scene_step(sphg_scene* s, uint32_t nsteps, sphg_step_stats* stats) {
    if (!s || !s->ctx) return fail(SPHG_EINVAL, "NULL scene");
    if (!s->ready) return fail(SPHG_ESTATE, "step before init");
    HIP_TRY(hipSetDevice(s->ctx->device));
    Batch& B = *s->self;
    int rc = self_tables(s);
    if (rc) return rc;
    rc = run_batch(s->ctx, B, nsteps, stats);
    if (rc) return rc;
    HIP_TRY(hipGetLastError());
    return SPHG_OK;
}

int sphg_scene_autotune(sphg_scene* s, uint32_t steps_per_trial, uint32_t rounds, sphg_tune_result* out) {
    if (!s || !s->ctx) return fail(SPHG_EINVAL, "NULL scene");
    if (!s->ready) return fail(SPHG_ESTATE, "autotune before init");
    HIP_TRY(hipSetDevice(s->ctx->device));
    int rc = self_tables(s);
    if (rc) return rc;
    rc = tune_batch(s->ctx, *s->self, steps_per_trial, rounds, out);
    if (rc) return rc;
    HIP_TRY(hipGetLastError());
    return SPHG_OK;
}

int sphg_batch_step(sphg_ctx* ctx, sphg_scene* const* scenes, uint32_t nscenes, uint32_t nsteps,
                    sphg_step_stats* stats) {
    if (!ctx || (!scenes && nscenes)) return fail(SPHG_EINVAL, "NULL argument");
    if (nscenes == 0) return SPHG_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    int rc = check_batch(ctx, scenes, nscenes);
    if (rc) return rc;
    Batch B;
    B.scenes.assign(scenes, scenes + nscenes);
    rc = upload_batch_tables(ctx, B);
    if (rc) return rc;
    rc = run_batch(ctx, B, nsteps, stats);
    if (rc) return rc;
    HIP_TRY(hipGetLastError());
    return SPHG_OK;
}

int sphg_batch_step_noalloc(sphg_ctx* ctx, sphg_scene* const* scenes, uint32_t nscenes, uint32_t nsteps,
                            sphg_step_stats* stats) {
    if (!ctx || (!scenes && nscenes)) return fail(SPHG_EINVAL, "NULL argument");
    if (nscenes == 0) return SPHG_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    int rc = check_batch(ctx, scenes, nscenes);
    if (rc) return rc;
    if (ctx->prof.on)
        return fail(SPHG_ESTATE, "kernel timing records HIP events; switch it off for the allocation-free step");
    if (!ctx->scratch) return fail(SPHG_ESTATE, "no scratch address for this context's device");
    // the tables' layout in the scratch array: two scene tables, then the
    // particle-chunk and cell-chunk tables, each at a 256-byte boundary
    size_t nP = 0, nC = 0;
    for (uint32_t i = 0; i < nscenes; ++i) {
        const DevScene& V = scenes[i]->view[scenes[i]->par];
        nP += (size_t)V.nChunks;
        nC += (size_t)V.nCellChunks;
    }
    auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t tab = up(sizeof(DevScene) * nscenes);
    const size_t offP = 2 * tab, offC = offP + up(sizeof(Chunk) * std::max<size_t>(nP, 1));
    const size_t need = offC + up(sizeof(Chunk) * std::max<size_t>(nC, 1));
    if (need > (size_t)SPHG_NOALLOC_SCRATCH_BYTES)
        return fail(SPHG_EINVAL, "the batch needs " + std::to_string(need) + " bytes of tables (" +
                                     std::to_string(nscenes) + " scenes, " + std::to_string(nP) + " particle chunks, " +
                                     std::to_string(nC) + " cell chunks), more than the " +
                                     std::to_string((size_t)SPHG_NOALLOC_SCRATCH_BYTES) +
                                     " bytes the allocation-free step holds; use sphg_batch_step");
    Batch B;
    B.ownTables = false;
    B.scenes.assign(scenes, scenes + nscenes);
    B.dT[0] = (DevScene*)(ctx->scratch);
    B.dT[1] = (DevScene*)(ctx->scratch + tab);
    B.dP = (Chunk*)(ctx->scratch + offP);
    B.dC = (Chunk*)(ctx->scratch + offC);
    B.tab0 = B.dT[0];
    rc = upload_batch_tables(ctx, B);
    if (rc) return rc;
    rc = run_batch(ctx, B, nsteps, stats, false);
    if (rc) return rc;
    HIP_TRY(hipGetLastError());
    return SPHG_OK;
}

int sphg_batch_create(sphg_ctx* ctx, sphg_scene* const* scenes, uint32_t nscenes, sphg_batch** out) {
    if (!ctx || !out || !scenes || nscenes == 0) return fail(SPHG_EINVAL, "NULL argument or empty batch");
    *out = nullptr;
    HIP_TRY(hipSetDevice(ctx->device));
    int rc = check_batch(ctx, scenes, nscenes);
    if (rc) return rc;
    sphg_batch* b = new sphg_batch();
    b->ctx = ctx;
    b->B.scenes.assign(scenes, scenes + nscenes);
    b->B.graphs = true;
    rc = upload_batch_tables(ctx, b->B);
    if (rc) {
        delete b;
        return rc;
    }
    *out = b;
    return SPHG_OK;
}

// A scene stepped elsewhere since the batch's tables were built has moved
// its double buffers: rebuild the tables from the scenes' current parity.
static int batch_tables_current(sphg_batch* b) {
    Batch& B = b->B;
    bool stale = false;
    for (size_t i = 0; i < B.scenes.size(); ++i)
        stale |= (B.parity[i] != B.scenes[i]->par) || (B.nAt[i] != B.scenes[i]->n);
    return stale ? upload_batch_tables(b->ctx, B) : SPHG_OK;
}

int sphg_batch_run(sphg_batch* b, uint32_t nsteps, sphg_step_stats* stats) {
    if (!b) return fail(SPHG_EINVAL, "NULL batch");
    sphg_ctx* ctx = b->ctx;
    HIP_TRY(hipSetDevice(ctx->device));
    int rc = batch_tables_current(b);
    if (rc) return rc;
    rc = run_batch(ctx, b->B, nsteps, stats);
    if (rc) return rc;
    HIP_TRY(hipGetLastError());
    return SPHG_OK;
}

int sphg_batch_autotune(sphg_batch* b, uint32_t steps_per_trial, uint32_t rounds, sphg_tune_result* out) {
    if (!b) return fail(SPHG_EINVAL, "NULL batch");
    HIP_TRY(hipSetDevice(b->ctx->device));
    int rc = batch_tables_current(b);
    if (rc) return rc;
    rc = tune_batch(b->ctx, b->B, steps_per_trial, rounds, out);
    if (rc) return rc;
    HIP_TRY(hipGetLastError());
    return SPHG_OK;
}

int sphg_batch_destroy(sphg_batch* b) {
    if (!b) return SPHG_OK;
    (void)hipSetDevice(b->ctx->device);
    delete b;
    return SPHG_OK;
}

int sphg_scene_plan_bytes(sphg_scene* s, uint64_t* bytes) {
    if (!s || !bytes) return fail(SPHG_EINVAL, "NULL argument");
    if (s->ready) {
        *bytes = s->memBytes;
        return SPHG_OK;
    }
    if (s->cap == 0) return fail(SPHG_ESTATE, "no fluid particles and no emitters");
    Plan P;
    int rc = make_plan(s, P);
    if (rc) return rc;
    *bytes = P.total;
    return SPHG_OK;
}

int sphg_scene_get(sphg_scene* s, int32_t field, void* out) {
    if (!s || !out) return fail(SPHG_EINVAL, "NULL argument");
    if (!s->ready) return fail(SPHG_ESTATE, "get before init");
    sphg_ctx* ctx = s->ctx;
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const DevScene& V = s->view[s->par];
    const int n = (int)s->n;
    size_t bytes = 0;
    // DFSPH forms a, the normals and rho* before the step's grid rebuild:
    // they are in the order of the ids before the rebuild's sort
    const uint32_t* pre = (V.method == DFSPH && s->stepped) ? s->view[1 - s->par].id : V.id;
    switch (field) {
        case SPHG_FIELD_X: launch_unsort_f4(st, V.x, V.id, s->tmp, n, 3); bytes = 12ull * n; break;
        case SPHG_FIELD_V: launch_unsort_f4(st, V.v, V.id, s->tmp, n, 3); bytes = 12ull * n; break;
        case SPHG_FIELD_A: launch_unsort_f4(st, V.a, pre, s->tmp, n, 3); bytes = 12ull * n; break;
        case SPHG_FIELD_RHO: launch_unsort_f1(st, V.drho, V.id, s->tmp, n, V.rho0D); bytes = 4ull * n; break;
        case SPHG_FIELD_DRHO: launch_unsort_f1(st, V.drho, V.id, s->tmp, n); bytes = 4ull * n; break;
        case SPHG_FIELD_DRHOSTAR:
            if (V.method != DFSPH) return fail(SPHG_EINVAL, "DFSPH field on another method");
            launch_unsort_f1(st, V.drhoS, pre, s->tmp, n); bytes = 4ull * n; break;
        case SPHG_FIELD_P:
            if (!V.p) return fail(SPHG_EINVAL, "no pressure field for this method");
            launch_unsort_f1(st, V.p, V.id, s->tmp, n); bytes = 4ull * n; break;
        case SPHG_FIELD_ALPHA:
        case SPHG_FIELD_RHOSTAR:
        case SPHG_FIELD_DIVERR:
        case SPHG_FIELD_K:
        case SPHG_FIELD_KV: {
            if (V.method != DFSPH) return fail(SPHG_EINVAL, "DFSPH field on another method");
            const float* src = field == SPHG_FIELD_ALPHA ? V.alpha : field == SPHG_FIELD_RHOSTAR ? V.drhoS
                             : field == SPHG_FIELD_DIVERR ? V.divErr : field == SPHG_FIELD_K ? V.k : V.kv;
            launch_unsort_f1(st, src, field == SPHG_FIELD_RHOSTAR ? pre : V.id, s->tmp, n,
                             field == SPHG_FIELD_RHOSTAR ? V.rho0D : 0.0);
            bytes = 4ull * n;
            break;
        }
        case SPHG_FIELD_NORMAL:
            if (!V.nrm) return fail(SPHG_EINVAL, "no surface tension in this scene");
            launch_unsort_f4(st, V.nrm, pre, s->tmp, n, 3); bytes = 12ull * n; break;
        case SPHG_FIELD_PACC:
            if (!V.pacc) return fail(SPHG_EINVAL, "PCISPH field on another method");
            launch_unsort_f4(st, V.pacc, V.id, s->tmp, n, 3); bytes = 12ull * n; break;
        case SPHG_FIELD_KEYS: {
            int rc = self_tables(s);
            if (rc) return rc;
            launch_keys3(st, s->self->dT[0], 0, n, (int*)s->tmp);
            bytes = 12ull * n;
            break;
        }
        case SPHG_FIELD_NSH: {
            if (!V.nsh) return fail(SPHG_EINVAL, "shell lists belong to PCISPH scenes");
            std::vector<uint16_t> c16(n);
            std::vector<uint32_t> ids(n);
            HIP_TRY(hipMemcpyAsync(ids.data(), V.id, 4ull * n, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipMemcpyAsync(c16.data(), V.nsh, 2ull * n, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
            int32_t* o = (int32_t*)out;
            for (int i = 0; i < n; ++i) o[ids[i]] = c16[i] == 0xFFFF ? -1 : (int32_t)c16[i];
            return SPHG_OK;
        }
        case SPHG_FIELD_NNB:
        case SPHG_FIELD_NGRAD: {
            std::vector<uint16_t> c16;
            std::vector<uint8_t> c8;
            std::vector<uint32_t> ids(n);
            HIP_TRY(hipMemcpyAsync(ids.data(), V.id, 4ull * n, hipMemcpyDeviceToHost, st));
            if (field == SPHG_FIELD_NNB) {
                c16.resize(n);
                HIP_TRY(hipMemcpyAsync(c16.data(), V.nnb, 2ull * n, hipMemcpyDeviceToHost, st));
            } else {
                c8.resize(n);
                HIP_TRY(hipMemcpyAsync(c8.data(), V.ngrad, 1ull * n, hipMemcpyDeviceToHost, st));
            }
            HIP_TRY(hipStreamSynchronize(st));
            int32_t* o = (int32_t*)out;
            for (int i = 0; i < n; ++i) o[ids[i]] = field == SPHG_FIELD_NNB ? (int32_t)c16[i] : (int32_t)c8[i];
            return SPHG_OK;
        }
        case SPHG_FIELD_PBNORMAL:
            if (!V.pbOn) return fail(SPHG_EINVAL, "fluid normals belong to boundaryMethod 1 (PCISPH)");
            launch_unsort_f4(st, V.pbn, V.id, s->tmp, n, 3); bytes = 12ull * n; break;
        case SPHG_FIELD_BDENSITY:
        case SPHG_FIELD_BPRESSURE: {
            if (!V.pbOn) return fail(SPHG_EINVAL, "boundary densities belong to boundaryMethod 1 (PCISPH)");
            std::vector<float> tmp(std::max<uint32_t>(s->nb, 1));
            HIP_TRY(hipMemcpyAsync(tmp.data(), field == SPHG_FIELD_BDENSITY ? V.bdens : V.bpres,
                                   sizeof(float) * s->nb, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
            float* o = (float*)out;
            for (uint32_t i = 0; i < s->nb; ++i) o[s->border[i]] = tmp[i];
            return SPHG_OK;
        }
        case SPHG_FIELD_BVOL: {
            std::vector<float4> xb(std::max<uint32_t>(s->nb, 1));
            HIP_TRY(hipMemcpyAsync(xb.data(), V.xb, sizeof(float4) * s->nb, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
            float* o = (float*)out;
            for (uint32_t i = 0; i < s->nb; ++i) o[s->border[i]] = xb[i].w;
            return SPHG_OK;
        }
        default: return fail(SPHG_EINVAL, "unknown field " + std::to_string(field));
    }
    HIP_TRY(hipMemcpyAsync(out, s->tmp, bytes, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    HIP_TRY(hipGetLastError());
    return SPHG_OK;
}

int sphg_scene_get_scalars(sphg_scene* s, sphg_scene_scalars* out) {
    if (!s || !out) return fail(SPHG_EINVAL, "NULL argument");
    if (!s->ready) return fail(SPHG_ESTATE, "get before init");
    HIP_TRY(hipSetDevice(s->ctx->device));
    Scalars v;
    HIP_TRY(hipMemcpyAsync(&v, s->ss, sizeof(Scalars), hipMemcpyDeviceToHost, s->ctx->stream));
    HIP_TRY(hipStreamSynchronize(s->ctx->stream));
    out->time = v.time;
    out->time_step = v.dt;
    out->max_vel = v.maxVel;
    out->cfl_time_step = v.cflTs;
    out->avg_density_error = v.avgErr;
    out->avg_density_error0 = v.avgErr0;
    out->avg_divergence_error = v.avgErrV;
    out->avg_divergence_error0 = v.avgErrV0;
    out->max_density_error = v.maxDensErr;
    out->iterations = v.it;
    out->iterations_v = v.itV;
    out->steps = v.steps;
    out->sum_iterations = v.sumIt;
    out->sum_iterations_v = v.sumItV;
    out->neighbor_overflow = v.overflow;
    out->pcisph_shell_steps = v.shellSteps;
    out->pcisph_walk_steps = v.walkSteps;
    return SPHG_OK;
}

int sphg_scene_get_sorted(sphg_scene* s, uint32_t* ids, int32_t* cells) {
    if (!s || !ids || !cells) return fail(SPHG_EINVAL, "NULL argument");
    if (!s->ready) return fail(SPHG_ESTATE, "get before init");
    const DevScene& V = s->view[s->par];
    hipStream_t st = s->ctx->stream;
    HIP_TRY(hipMemcpyAsync(ids, V.id, 4ull * s->n, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(cells, V.cell, 4ull * s->n, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return SPHG_OK;
}

int sphg_scene_get_neighbors(sphg_scene* s, uint32_t* counts, uint32_t* ids, uint32_t cap) {
    if (!s || !counts || !ids) return fail(SPHG_EINVAL, "NULL argument");
    if (!s->ready) return fail(SPHG_ESTATE, "get before init");
    const DevScene& V = s->view[s->par];
    if ((int)cap < V.maxNbr) return fail(SPHG_EINVAL, "cap below max_neighbors");
    const size_t n = s->n;
    const size_t npad = (size_t)V.nChunks * kBlock;
    std::vector<uint32_t> sid(n), lst(npad * (size_t)V.maxNbr);
    std::vector<uint16_t> c16(n);
    hipStream_t st = s->ctx->stream;
    HIP_TRY(hipMemcpyAsync(sid.data(), V.id, 4 * n, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(c16.data(), V.nnb, 2 * n, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(lst.data(), V.nbr, 4 * lst.size(), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (size_t i = 0; i < n; ++i) {
        const uint32_t me = sid[i];
        counts[me] = c16[i];
        // blocks of 4 slots, lane-major within a block (see slot_of)
        const size_t base = (i >> 6) * (size_t)V.maxNbr * kGroup + (i & 63) * 4;
        for (uint32_t q = 0; q < c16[i]; ++q)
            ids[(size_t)me * cap + q] = sid[lst[base + (size_t)(q >> 2) * kGroup * 4 + (q & 3)]];
    }
    return SPHG_OK;
}

int sphg_snapshot_create(sphg_scene* s, sphg_snapshot** out) {
    if (!s || !s->ctx || !out) return fail(SPHG_EINVAL, "NULL argument");
    if (!s->ready) return fail(SPHG_ESTATE, "snapshot before init");
    HIP_TRY(hipSetDevice(s->ctx->device));
    if (!s->ctx->copyStream) HIP_TRY(hipStreamCreateWithFlags(&s->ctx->copyStream, hipStreamNonBlocking));
    sphg_snapshot* sn = new sphg_snapshot();
    sn->scene = s;
    sn->capBytes = sizeof(float) * 3 * (size_t)std::max<uint32_t>(s->cap, 1);
    const size_t total = 2 * sn->capBytes + sizeof(Scalars);
    hipError_t e = hipMalloc(&sn->dStage, total);
    if (e == hipSuccess) e = hipHostMalloc(&sn->hStage, total, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&sn->ready, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&sn->done, hipEventDisableTiming);
    if (e != hipSuccess) {
        sn->release();
        delete sn;
        return fail(SPHG_ENOMEM, std::string("snapshot buffers: ") + hipGetErrorString(e));
    }
    s->snaps.push_back(sn);
    *out = sn;
    return SPHG_OK;
}

int sphg_snapshot_begin(sphg_snapshot* sn) {
    if (!sn) return fail(SPHG_EINVAL, "NULL snapshot");
    sphg_scene* s = sn->scene;
    if (!s || !s->ctx) return fail(SPHG_ESTATE, "the snapshot's scene was destroyed");
    sphg_ctx* ctx = s->ctx;
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    // the staging buffer is free once this snapshot's previous copy is done
    if (sn->pending) HIP_TRY(hipStreamWaitEvent(st, sn->done, 0));
    const DevScene& V = s->view[s->par];
    const int n = (int)s->n;
    char* d = (char*)sn->dStage;
    launch_unsort_f4(st, V.x, V.id, (float*)d, n, 3);
    launch_unsort_f4(st, V.v, V.id, (float*)(d + sn->capBytes), n, 3);
    HIP_TRY(hipMemcpyAsync(d + 2 * sn->capBytes, s->ss, sizeof(Scalars), hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipEventRecord(sn->ready, st));
    hipStream_t cs = ctx->copyStream;
    HIP_TRY(hipStreamWaitEvent(cs, sn->ready, 0));
    char* h = (char*)sn->hStage;
    const size_t b = sizeof(float) * 3 * (size_t)n;
    if (b) {
        HIP_TRY(hipMemcpyAsync(h, d, b, hipMemcpyDeviceToHost, cs));
        HIP_TRY(hipMemcpyAsync(h + sn->capBytes, d + sn->capBytes, b, hipMemcpyDeviceToHost, cs));
    }
    HIP_TRY(hipMemcpyAsync(h + 2 * sn->capBytes, d + 2 * sn->capBytes, sizeof(Scalars), hipMemcpyDeviceToHost, cs));
    HIP_TRY(hipEventRecord(sn->done, cs));
    HIP_TRY(hipGetLastError());
    sn->n = s->n;
    sn->pending = true;
    return SPHG_OK;
}

int sphg_snapshot_wait(sphg_snapshot* sn, const float** xyz, const float** vel, uint32_t* n, double* time,
                       double* time_step) {
    if (!sn) return fail(SPHG_EINVAL, "NULL snapshot");
    if (!sn->pending) return fail(SPHG_ESTATE, "snapshot never begun");
    if (!sn->done) return fail(SPHG_ESTATE, "the snapshot's scene was destroyed");
    HIP_TRY(hipEventSynchronize(sn->done));
    const char* h = (const char*)sn->hStage;
    Scalars sc;
    std::memcpy(&sc, h + 2 * sn->capBytes, sizeof(Scalars));
    if (xyz) *xyz = (const float*)h;
    if (vel) *vel = (const float*)(h + sn->capBytes);
    if (n) *n = sn->n;
    if (time) *time = sc.time;
    if (time_step) *time_step = sc.dt;
    return SPHG_OK;
}

int sphg_snapshot_destroy(sphg_snapshot* sn) {
    if (!sn) return SPHG_OK;
    if (sn->scene) {
        auto& v = sn->scene->snaps;
        v.erase(std::remove(v.begin(), v.end(), sn), v.end());
    }
    sn->release();
    delete sn;
    return SPHG_OK;
}

int sphg_scene_info(sphg_scene* s, uint32_t* n, uint32_t* nb, uint32_t* ncells, uint64_t* bytes) {
    if (!s) return fail(SPHG_EINVAL, "NULL scene");
    if (n) *n = s->n;
    if (nb) *nb = s->nb;
    if (ncells) *ncells = s->ready ? (uint32_t)s->view[0].ncells : 0;
    if (bytes) *bytes = s->memBytes;
    return SPHG_OK;
}

int sphg_ctx_kernel_timing(sphg_ctx* ctx, int32_t enable) {
    if (!ctx) return fail(SPHG_EINVAL, "ctx is NULL");
    HIP_TRY(hipSetDevice(ctx->device));
    ctx->prof.collect();
    ctx->prof.on = enable != 0;
    ctx->prof.mask = ~0ull;
    ctx->prof.stride = 1;
    if (enable) {
        for (int k = 0; k < K_NKINDS; ++k) {
            ctx->prof.count[k] = 0;
            ctx->prof.ms[k] = 0.0;
            ctx->prof.seen[k] = 0;
        }
    }
    return SPHG_OK;
}

int sphg_ctx_kernel_timing_stride(sphg_ctx* ctx, uint32_t stride) {
    if (!ctx || stride == 0) return fail(SPHG_EINVAL, "NULL ctx or zero stride");
    ctx->prof.stride = stride;
    return SPHG_OK;
}

int sphg_ctx_kernel_timing_only(sphg_ctx* ctx, const char* const* names, uint32_t count) {
    if (!ctx || (count && !names)) return fail(SPHG_EINVAL, "NULL argument");
    uint64_t mask = 0;
    for (uint32_t i = 0; i < count; ++i) {
        int k = 0;
        while (k < K_NKINDS && (!names[i] || std::strcmp(names[i], kKindNames[k]) != 0)) ++k;
        if (k == K_NKINDS)
            return fail(SPHG_EINVAL, std::string("unknown kernel name ") + (names[i] ? names[i] : "(null)"));
        mask |= 1ull << k;
    }
    const int rc = sphg_ctx_kernel_timing(ctx, 1);
    if (rc) return rc;
    ctx->prof.mask = mask;
    return SPHG_OK;
}

int sphg_ctx_kernel_time(sphg_ctx* ctx, const char* name, uint64_t* launches, double* total_ms) {
    if (!ctx || !name) return fail(SPHG_EINVAL, "NULL argument");
    ctx->prof.collect();
    for (int k = 0; k < K_NKINDS; ++k)
        if (std::strcmp(name, kKindNames[k]) == 0) {
            if (launches) *launches = ctx->prof.count[k];
            if (total_ms) *total_ms = ctx->prof.ms[k];
            return SPHG_OK;
        }
    return fail(SPHG_EINVAL, std::string("unknown kernel name ") + name);
}

// ------------------------------------------------------------- samplers

int64_t sphg_sample_box(const double* mn, const double* mx, double supportRadius, float* out, int64_t cap) {
    if (!mn || !mx || !(supportRadius > 0)) return -1;
    const double lx = std::fabs(mn[0] - mx[0]), ly = std::fabs(mn[1] - mx[1]), lz = std::fabs(mn[2] - mx[2]);
    const unsigned nx = (unsigned)std::ceil(lx / (0.5 * supportRadius));
    const unsigned ny = (unsigned)std::ceil(ly / (0.5 * supportRadius));
    const unsigned nz = (unsigned)std::ceil(lz / (0.5 * supportRadius));
    const double dx = lx / nx, dy = ly / ny, dz = lz / nz;
    int64_t c = 0;
    for (unsigned i = 0; i <= nx; i++)
        for (unsigned j = 0; j <= ny; j++)
            for (unsigned k = 0; k <= nz; k++)
                if (i == 0 || i == nx || j == 0 || j == ny || k == 0 || k == nz) {
                    if (out && c < cap) {
                        out[3 * c] = (float)(i * dx + mn[0]);
                        out[3 * c + 1] = (float)(j * dy + mn[1]);
                        out[3 * c + 2] = (float)(k * dz + mn[2]);
                    }
                    ++c;
                }
    return c;
}

int64_t sphg_sample_sphere(const double* o, double radius, double pr, float* out, int64_t cap) {
    if (!o || !(radius > 0) || !(pr > 0)) return -1;
    const double sep = 1.5;
    const unsigned num = (unsigned)std::floor(4.0 * M_PI * radius * radius / (M_PI * pr * pr) * sep);
    const double a = 4.0 * M_PI / num;
    const double dd = std::sqrt(a);
    const int mt = int(std::round(M_PI / dd));
    const double dth = M_PI / mt;
    const double dph = a / dth;
    int64_t c = 0;
    for (int m = 0; m < mt; ++m) {
        const double th = M_PI * (m + 0.5) / mt;
        const int mp = int(std::round(2.0 * M_PI * std::sin(th) / dph));
        for (int q = 0; q < mp; ++q) {
            const double ph = 2.0 * M_PI * q / mp;
            if (out && c < cap) {
                out[3 * c] = (float)(radius * (std::sin(th) * std::cos(ph)) + o[0]);
                out[3 * c + 1] = (float)(radius * (std::sin(th) * std::sin(ph)) + o[1]);
                out[3 * c + 2] = (float)(radius * std::cos(th) + o[2]);
            }
            ++c;
        }
    }
    return c;
}

// PCISPHBoundaryModel::sampleCube + addCube (src/PCISPHBoundaryModel.cpp:
// 7-27, 60-189): the box surface at ceil(l / 0.45h) intervals per axis.
// A point's normal is the mean of the inward normals of the faces it lies on
// (+x at the minimum x face, -x at the maximum, and so on), glm::normalize'd;
// the face normals are axis unit vectors, so their sums are exact in any
// order.  Inverted keeps these normals, otherwise they are negated.
int64_t sphg_sample_box_normals(const double* mn, const double* mx, double supportRadius, int32_t inverted, float* out,
                                float* nrm, int64_t cap) {
    if (!mn || !mx || !(supportRadius > 0)) return -1;
    const double l[3] = {std::fabs(mn[0] - mx[0]), std::fabs(mn[1] - mx[1]), std::fabs(mn[2] - mx[2])};
    if (!(l[0] > 0 && l[1] > 0 && l[2] > 0)) return -1;
    unsigned nn[3];
    double dd[3];
    for (int a = 0; a < 3; ++a) {
        nn[a] = (unsigned)std::ceil(l[a] / (0.45 * supportRadius));
        dd[a] = l[a] / nn[a];
    }
    const double inv = inverted ? 1.0 : -1.0;
    int64_t c = 0;
    for (unsigned i = 0; i <= nn[0]; i++)
        for (unsigned j = 0; j <= nn[1]; j++)
            for (unsigned k = 0; k <= nn[2]; k++) {
                const unsigned ix[3] = {i, j, k};
                double v[3] = {0, 0, 0};
                int faces = 0;
                for (int a = 0; a < 3; ++a) {
                    if (ix[a] == 0) { v[a] = 1.0; ++faces; }
                    else if (ix[a] == nn[a]) { v[a] = -1.0; ++faces; }
                }
                if (!faces) continue;
                if (out && c < cap) {
                    for (int a = 0; a < 3; ++a) out[3 * c + a] = (float)(ix[a] * dd[a] + mn[a]);
                    if (faces == 2) for (double& e : v) e *= 0.5;
                    if (faces == 3) for (double& e : v) e /= 3.0;
                    const double f = 1.0 / std::sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
                    if (nrm)
                        for (int a = 0; a < 3; ++a) nrm[3 * c + a] = (float)(inv * (v[a] * f));
                }
                ++c;
            }
    return c;
}

// PCISPHBoundaryModel::addSphere (src/PCISPHBoundaryModel.cpp:29-52): the
// sphere points of sphg_sample_sphere with normals inv (center - point)
// (unnormalised), inv = 1 if inverted else -1.
int64_t sphg_sample_sphere_normals(const double* o, double radius, double pr, int32_t inverted, float* out,
                                   float* nrm, int64_t cap) {
    if (!o || !(radius > 0) || !(pr > 0)) return -1;
    const double sep = 1.5;
    const unsigned num = (unsigned)std::floor(4.0 * M_PI * radius * radius / (M_PI * pr * pr) * sep);
    const double a = 4.0 * M_PI / num;
    const double dd = std::sqrt(a);
    const int mt = int(std::round(M_PI / dd));
    const double dth = M_PI / mt;
    const double dph = a / dth;
    const double inv = inverted ? 1.0 : -1.0;
    int64_t c = 0;
    for (int m = 0; m < mt; ++m) {
        const double th = M_PI * (m + 0.5) / mt;
        const int mp = int(std::round(2.0 * M_PI * std::sin(th) / dph));
        for (int q = 0; q < mp; ++q) {
            const double ph = 2.0 * M_PI * q / mp;
            if (out && c < cap) {
                const double p[3] = {radius * (std::sin(th) * std::cos(ph)) + o[0],
                                     radius * (std::sin(th) * std::sin(ph)) + o[1], radius * std::cos(th) + o[2]};
                for (int e = 0; e < 3; ++e) {
                    out[3 * c + e] = (float)p[e];
                    if (nrm) nrm[3 * c + e] = (float)(inv * (o[e] - p[e]));
                }
            }
            ++c;
        }
    }
    return c;
}

static inline uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static inline double unit01(uint64_t& s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

int64_t sphg_fluid_block(const double* mn, const double* mx, double pr, double jitter, uint64_t seed, float* out,
                         int64_t cap) {
    if (!mn || !mx || !(pr > 0)) return -1;
    const double dist = 2.0 * pr;
    const int px = (int)std::floor((mx[0] - mn[0]) / dist + 1e-5);
    const int py = (int)std::floor((mx[1] - mn[1]) / dist + 1e-5);
    const int pz = (int)std::floor((mx[2] - mn[2]) / dist + 1e-5);
    if (px <= 0 || py <= 0 || pz <= 0) return 0;
    uint64_t st = seed;
    int64_t c = 0;
    for (int i = 0; i < px; ++i)
        for (int j = 0; j < py; ++j)
            for (int k = 0; k < pz; ++k) {
                double p[3] = {(double)i * dist + (mn[0] + pr), (double)j * dist + (mn[1] + pr),
                               (double)k * dist + (mn[2] + pr)};
                if (jitter != 0.0)
                    for (int a = 0; a < 3; ++a) p[a] += (2.0 * unit01(st) - 1.0) * jitter * dist;
                if (out && c < cap)
                    for (int a = 0; a < 3; ++a) out[3 * c + a] = (float)p[a];
                ++c;
            }
    return c;
}

int64_t sphg_fluid_spheres(const double* mn, const double* mx, double pr, const double* centers, const double* radii,
                           int32_t ns, float* out, int64_t cap) {
    if (!mn || !mx || !(pr > 0) || (ns > 0 && (!centers || !radii))) return -1;
    const double dist = 2.0 * pr;
    const int px = (int)std::floor((mx[0] - mn[0]) / dist + 1e-5);
    const int py = (int)std::floor((mx[1] - mn[1]) / dist + 1e-5);
    const int pz = (int)std::floor((mx[2] - mn[2]) / dist + 1e-5);
    int64_t c = 0;
    for (int i = 0; i < px; ++i)
        for (int j = 0; j < py; ++j)
            for (int k = 0; k < pz; ++k) {
                const double p[3] = {(double)i * dist + (mn[0] + pr), (double)j * dist + (mn[1] + pr),
                                     (double)k * dist + (mn[2] + pr)};
                bool in = false;
                for (int s = 0; s < ns && !in; ++s) {
                    const double dx = p[0] - centers[3 * s], dy = p[1] - centers[3 * s + 1], dz = p[2] - centers[3 * s + 2];
                    in = dx * dx + dy * dy + dz * dz <= radii[s] * radii[s];
                }
                if (!in) continue;
                if (out && c < cap)
                    for (int a = 0; a < 3; ++a) out[3 * c + a] = (float)p[a];
                ++c;
            }
    return c;
}

}  // extern "C"
