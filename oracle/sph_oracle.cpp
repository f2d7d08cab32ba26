// sph_oracle.cpp -- FP64 CPU restatement of the reference SPH solver path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the oracle the HIP solver in
// sph_amd/csrc is tested against: only tests/, __graft_entry__.smoke() and
// bench.py's CPU baseline leg may load it.  The product path never calls it.
//
// It restates, in double precision and with the reference's own data
// structures (an unordered_map cell hash whose buckets hold per-cell
// candidate lists, AoS double state, OpenMP parallel-for over particles),
// the algorithm of josemanuel28v/SPH:
//   grid            src/HashTable.h:79-169, src/SPHSolver.cpp:172-221
//   kernels         src/CubicSpline.h:18-73, src/Poly6.h:21-57,
//                   src/Spiky.h:20-55, src/Cohesion.h:20-40,
//                   src/Adhesion.h:18-35
//   densities       src/SPHSolver.cpp:44-139
//   Akinci volumes  src/AkinciBoundayModel.cpp:11-40
//   samplers        src/AkinciBoundayModel.cpp:100-164,
//                   src/Simulation.cpp:512-539
//   WCSPH           src/WCSPHSolver.cpp:14-188, src/SPHSolver.cpp:141-170
//   PCISPH          src/PCISPHSolver.cpp:12-474
//   DFSPH           src/DFSPHSolver.cpp:10-870
//   forces          src/XSPHViscosity.h:32-69,
//                   src/SurfaceTensionAkinci.h:34-98,
//                   src/AdhesionForce.h:30-62
//   time            src/TimeManager.h:55, src/Simulation.cpp:134-146
//
// Vector arithmetic follows glm semantics: dot(a,b) = (a*b).x + (a*b).y +
// (a*b).z summed left to right, length = sqrt(dot(v,v)).  Build with
// -ffp-contract=off for the parity variant so every operation is a single
// IEEE-754 rounding (oracle/Makefile).
//
// One fluid model and one Akinci boundary model per scene: the reference's
// fluid neighbor loop ignores the model id (src/Simulation.h:24-29) and its
// boundary-volume sum indexes the own model's arrays with every boundary
// model's ids (src/AkinciBoundayModel.cpp:32), so scenes with more models are
// not meaningful there either.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdio>
#include <string>
#include <unordered_map>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// Planted mistakes.  oracle_plant(id) switches one deliberate error on, so
// the tests can show that each of their checks fails on the mistake it is
// meant to catch; 0 (the default) is the faithful restatement.
enum Plant {
    PL_NONE = 0,
    PL_KAPPA_X2,        // kappa twice as large (src/DFSPHSolver.cpp:346, :531)
    PL_KAPPA_DT,        // density kappa divided by dt instead of dt^2 (:346)
    PL_KAPPA_RHO0,      // kappa divided by rho0 instead of rho_i
    PL_K_NO_ALPHA,      // k += rho* - rho0, without alpha (:349, :534)
    PL_SWEEP_SIGN,      // v += instead of v -= in the velocity sweep (:353-394)
    PL_KV_DT2,          // kappa_v divided by dt^2 instead of dt (:531)
    PL_DIV_LOOP_MAXV,   // divergence loop bounded by maxIterationsV (:512 uses maxIterations)
    PL_ALPHA_NO_SUMSQ,  // alpha without the sum of squared gradients (:180-186)
    PL_DT_ABS_V,        // |v| instead of sqrt(|v|) in the CFL rule (:833, :848)
    PL_DT_NO_G,         // the CFL speed without (a + g) dt (:848)
    PL_DT_099_GE3,      // the 0.99 branch after 3 or more iterations (:859 tests > 3)
    PL_WC_P_OVER_RHO,   // p/rho instead of p/rho^2 (src/WCSPHSolver.cpp:159)
    PL_WC_PJ_AS_PI,     // p_i/rho_i^2 for both particles of a pair (:159)
    PL_WC_BOUNDARY_SYM, // 2 p_i/rho_i^2 at the boundary (:184)
    PL_XSPH_NO_DT,      // XSPH not divided by dt (src/XSPHViscosity.h:55)
    PL_XSPH_RHO_I,      // XSPH weighted by m/rho_i instead of m/rho_j (:55)
    PL_PCI_NO_CLAMP,    // p += delta (rho - rho0) without max(0, .) (src/PCISPHSolver.cpp:262)
    PL_PCI_CUBIC,       // cubic spline instead of Spiky gradient (:313-353)
    PL_PCI_P_OVER_RHO,  // p/rho instead of p/rho^2 in the pressure acceleration (:325)
    PL_ST_NPLUS,        // (n_i + n_j) in the curvature term (src/SurfaceTensionAkinci.h:90)
    PL_ST_K_HALF,       // K_ij = rho0 / (rho_i + rho_j) (:86)
    PL_ADH_COEF,        // adhesion scaled by 0.007/h^3.25 (src/Adhesion.h:21 computes it, :30-31 omit it)
    PL_ADH_NO_RHO0,     // adhesion without rho0 (src/AdhesionForce.h:57)
    PL_COUNT
};
int g_plant = PL_NONE;
inline bool planted(int p) { return g_plant == p; }

struct V3 {
    double x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(double a, double b, double c) : x(a), y(b), z(c) {}
};
inline V3 operator+(V3 a, V3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator-(V3 a) { return V3(-a.x, -a.y, -a.z); }
inline V3 operator*(V3 a, double s) { return V3(a.x * s, a.y * s, a.z * s); }
inline V3 operator*(double s, V3 a) { return V3(s * a.x, s * a.y, s * a.z); }
inline V3 operator/(V3 a, double s) { return V3(a.x / s, a.y / s, a.z / s); }
inline V3& operator+=(V3& a, V3 b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
inline V3& operator-=(V3& a, V3 b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; return a; }
inline V3& operator*=(V3& a, double s) { a.x *= s; a.y *= s; a.z *= s; return a; }
// glm::dot: component products summed left to right.
inline double dot(V3 a, V3 b) { double tx = a.x * b.x, ty = a.y * b.y, tz = a.z * b.z; return tx + ty + tz; }
inline double length(V3 a) { return std::sqrt(dot(a, a)); }

struct Key {
    int x, y, z;
    bool operator==(const Key& o) const { return x == o.x && y == o.y && z == o.z; }
};
// Any hash works: the bucket contents and candidate order do not depend on it.
struct KeyHash {
    size_t operator()(const Key& k) const {
        uint64_t h = (uint64_t)(uint32_t)k.x * 0x9E3779B97F4A7C15ull;
        h ^= (uint64_t)(uint32_t)k.y * 0xC2B2AE3D27D4EB4Full + (h << 6) + (h >> 2);
        h ^= (uint64_t)(uint32_t)k.z * 0x165667B19E3779F9ull + (h << 6) + (h >> 2);
        return (size_t)h;
    }
};
// (id, point-set id) pairs, as the reference's bucket entries (src/HashTable.h:39-43).
struct PointInfo { uint32_t id, set; };
struct Bucket {
    std::vector<PointInfo> points, neighbors, bpoints, bneighbors;
};

// key = floor(x / h) per component (src/HashTable.h:81,96).
inline Key key_of(V3 p, double h) {
    return Key{(int)std::floor(p.x / h), (int)std::floor(p.y / h), (int)std::floor(p.z / h)};
}

// ---------------------------------------------------------------- kernels
struct Kernels {
    double h;
    // CubicSpline (src/CubicSpline.h:18-27)
    double ck, cl;
    // Poly6 (src/Poly6.h:21-28)
    double h2, p6v, p6g;
    // Spiky (src/Spiky.h:20-27)
    double sp_v, sp_g;
    // Cohesion (src/Cohesion.h:20-27)
    double co1, co2;
    // Adhesion coefficient, computed but unused by W (src/Adhesion.h:18-22)
    double adc;

    void set(double sr) {
        h = sr;
        const double pi = M_PI;
        const double h3 = sr * sr * sr;
        ck = 8.0 / (pi * h3);
        cl = 48.0 / (pi * h3);
        h2 = sr * sr;
        const double h9 = sr * sr * sr * sr * sr * sr * sr * sr * sr;
        p6v = 315.0 / (64.0 * M_PI * h9);
        p6g = 945.0 / (32.0 * M_PI * h9);
        const double h6 = sr * sr * sr * sr * sr * sr;
        sp_v = 15.0 / (M_PI * h6);
        sp_g = 45.0 / (M_PI * h6);
        co1 = 32.0 / (M_PI * h9);
        co2 = h6 / 64.0;
        adc = 0.007 / std::pow(sr, 3.25);
    }
    // src/CubicSpline.h:29-48
    double W(double r) const {
        double res = 0.0;
        const double q = r / h;
        if (q <= 1.0) {
            if (q <= 0.5) {
                const double q2 = q * q;
                const double q3 = q2 * q;
                res = ck * (6.0 * q3 - 6.0 * q2 + 1.0);
            } else {
                res = ck * (2.0 * std::pow(1.0 - q, 3.0));
            }
        }
        return res;
    }
    double W(V3 r) const { return W(length(r)); }
    // src/CubicSpline.h:55-73
    V3 gradW(V3 r) const {
        V3 res;
        const double rl = length(r);
        const double q = rl / h;
        if ((rl > 1.0e-5) && (q <= 1.0)) {
            const V3 gradq = r * (1.0 / (rl * h));
            if (q <= 0.5) {
                res = cl * q * (3.0 * q - 2.0) * gradq;
            } else {
                const double factor = 1.0 - q;
                res = cl * (-factor * factor) * gradq;
            }
        }
        return res;
    }
    // src/Poly6.h:31-42
    double poly6W(V3 r) const {
        const double mr = length(r);
        const double mr2 = mr * mr;
        if (mr2 <= h2) {
            const double diff = h2 - mr2;
            return p6v * diff * diff * diff;
        }
        return 0.0;
    }
    // src/Poly6.h:44-57
    V3 poly6Grad(V3 r) const {
        double mr = length(r);
        double mr2 = mr * mr;
        if (mr2 <= h2) {
            double diff = h2 - mr2;
            return -diff * diff * p6g * r;
        }
        return V3();
    }
    // src/Spiky.h:44-55
    V3 spikyGrad(V3 r) const {
        double mr = length(r);
        if (mr != 0 && mr <= h) {
            double diff = h - mr;
            return -sp_g * diff * diff / mr * r;
        }
        return V3();
    }
    // src/Cohesion.h:29-40 (note: the h^6/64 term is subtracted after the
    // 32/(pi h^9) factor has been applied, as written there)
    double cohesion(V3 r) const {
        double value = 0;
        double mr = length(r);
        if ((2.0 * mr > h) && (mr <= h))
            value = co1 * std::pow(h - mr, 3.0) * std::pow(mr, 3.0);
        else if ((mr > 0) && (2.0 * mr <= h))
            value = co1 * 2.0 * std::pow(h - mr, 3.0) * std::pow(mr, 3.0) - co2;
        return value;
    }
    // src/Adhesion.h:24-35 (no 0.007/h^3.25 factor in the returned value)
    double adhesion(V3 r) const {
        double mr = length(r);
        if ((2.0 * mr > h) && (mr <= h)) {
            double value = -4.0 * mr * mr / h + 6 * mr - 2 * h;
            return std::pow(value, 1.0 / 4.0);
        }
        return 0.0;
    }
};

}  // namespace

// ------------------------------------------------------------------ C API
extern "C" {

// Scene parameters; field meaning follows the reference scene JSON
// (src/SceneLoader.cpp:38-141) and solver setters (src/Simulation.cpp:342-424).
typedef struct oracle_desc {
    int method;              // 0 WCSPH, 1 PCISPH, 2 DFSPH (src/Simulation.cpp:21-23)
    double particle_radius;
    double density0;
    double gravity[3];
    double time_step, min_time_step, max_time_step, fps;
    double cfl_factor;
    double stiffness, gamma;
    double eta, eta_v;
    unsigned min_iterations, max_iterations, min_iterations_v, max_iterations_v;
    int viscosity_method;    // -1 none, 0 standard, 1 artificial, 2 XSPH
    double viscosity, boundary_viscosity;
    int surface_tension_method;  // -1 none, 3 Akinci
    double surface_tension;
    int adhesion_method;     // -1 none, 4 Akinci
    double adhesion;
    int cube_enabled;        // CubeBoundaryModel (src/CubeBoundaryModel.cpp:5-94)
    double cube_min[3], cube_max[3];
    double cube_normal_fct, cube_tangential_fct;
    int boundary_method;     // 0 cube, 1 PCISPH, 2 Akinci (src/Simulation.cpp:25-27)
    double pb_normal_fct, pb_tangential_fct;   // PCISPHBoundaryModel (src/PCISPHBoundaryModel.h:21-22)
} oracle_desc;

enum {
    OF_X = 0, OF_V = 1, OF_A = 2, OF_RHO = 3, OF_P = 4, OF_ALPHA = 5, OF_RHOSTAR = 6,
    OF_DIVERR = 7, OF_K = 8, OF_KV = 9, OF_BVOL = 10, OF_NORMAL = 11, OF_PACC = 12,
    OF_ACC_STEP = 13, OF_BX = 14, OF_PBN = 15, OF_BDENS = 16, OF_BPRESS = 17, OF_PREDR = 18
};

// One emitter (src/Emitter.h:9-55, FluidModel::addEmitter
// src/FluidModel.cpp:124-143).  rotation is the quaternion (w, x, y, z) of
// the scene file's "rotation" array (src/SceneLoader.cpp:128-131).
typedef struct oracle_emitter {
    unsigned type;           // 0 square, 1 circle (src/Emitter.cpp:7-8)
    unsigned num_particles;
    double position[3];
    double velocity;
    double rotation[4];
    double start_time;
    double width, height, spacing;
} oracle_emitter;

}  // extern "C"

namespace {

inline V3 cross(V3 a, V3 b) { return V3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y); }

// Quaternion times vector as glm evaluates it: with u = (q.x, q.y, q.z),
// v + ((u x v) w + u x (u x v)) 2 (a rotation when |q| = 1).
inline V3 quatRotate(const double q[4], V3 v) {
    const V3 u(q[1], q[2], q[3]);
    const V3 uv = cross(u, v);
    const V3 uuv = cross(u, uv);
    return v + ((uv * q[0]) + uuv) * 2.0;
}

struct Emitter {
    oracle_emitter e;
    unsigned remaining = 0;      // numParticles: what is still to emit
    double nextTimeEmit = 0;
    std::vector<V3> group;

    // src/Emitter.cpp:10-29
    void setup(const oracle_emitter& ee, double particleRadius) {
        e = ee;
        remaining = e.num_particles;
        nextTimeEmit = e.start_time;
        if (e.type == 1) buildCircleGroup(particleRadius);
        else if (e.type == 0) buildSquareGroup(particleRadius);
    }
    // src/Emitter.cpp:31-55: the centre, then rings of radius (l+1) dist
    // holding floor(2 pi radius / dist) particles at angles i 2pi/nParts.
    void buildCircleGroup(double particleRadius) {
        const double dist = 2.0 * particleRadius;
        const double lv = std::floor((e.width - 0.5 * dist) / dist);
        const unsigned numLevels = lv > 0 ? (unsigned)lv : 0u;
        group.clear();
        group.push_back(V3(0.0, 0.0, 0.0));
        for (unsigned l = 0; l < numLevels; ++l) {
            const double radius = (l + 1) * dist;
            const double len = 2.0 * M_PI * radius;
            const unsigned nParts = (unsigned)std::floor(len / dist);
            const double inc = 2 * M_PI / nParts;
            for (unsigned i = 0; i < nParts; ++i) {
                const double angle = i * inc;
                group.push_back(V3(radius * std::sin(angle), radius * std::cos(angle), 0.0));
            }
        }
    }
    // src/Emitter.cpp:57-75: numX x numY lattice at spacing dist, centred,
    // x index outer, y inner, decreasing from +half size.
    void buildSquareGroup(double particleRadius) {
        const double dist = particleRadius * 2.0;
        const unsigned numX = (unsigned)std::floor(e.width / dist);
        const unsigned numY = (unsigned)std::floor(e.height / dist);
        const double realWidth = (double)(numX - 1) * dist;
        const double realHeight = (double)(numY - 1) * dist;
        group.clear();
        for (unsigned i = 0; i < numX; ++i)
            for (unsigned j = 0; j < numY; ++j)
                group.push_back(V3(realWidth * 0.5 - i * dist, realHeight * 0.5 - j * dist, 0.0));
    }
};

struct Oracle {
    oracle_desc d;
    Kernels K;
    double r = 0, h = 0, m = 0, rho0 = 1000;
    V3 g;
    // TimeManager (src/TimeManager.h:41-66)
    double time = 0, ts = 1e-4, minTs = 0, maxTs = 0, fps = 0;

    std::vector<V3> x, v, a, accStep;
    std::vector<double> rho, p;
    std::vector<V3> xb;
    std::vector<double> vb;

    std::unordered_map<Key, Bucket, KeyHash> table;

    // DFSPH state (src/DFSPHSolver.h:12-31)
    std::vector<double> alpha, divErr, rhoStar, k, kv;
    double avgDensError = 0, avgDivError = 0, avgDensError0 = 0, avgDivError0 = 0;
    unsigned iterations = 0, iterationsV = 0, steps = 0, sumIt = 0, sumItV = 0;
    bool f32Positions = false;   // round DFSPH's updated positions to FP32 (test emulation of FP32 storage)
    double eps = 1e-5;   // src/DFSPHSolver.cpp:26 (oracle_set_eps moves it for the tests)
    double maxVel = 0, lastCfl = 0;

    // PCISPH state (src/PCISPHSolver.h:12-17)
    std::vector<V3> predR, predV, pacc;
    double delta = 0, maxDensityError = 0;

    // Akinci surface tension normals (src/SurfaceTensionAkinci.h:19)
    std::vector<V3> normal;

    bool initialised = false;
    std::vector<Emitter> emitters;

    // PCISPH boundary method (src/PCISPHBoundaryModel.h:13-22): boundary
    // normals, densities and pressures; the fluid's normal field
    // (FluidModel::n, src/FluidModel.h:22), written by the corrections.
    bool pbm = false;   // PCISPH boundary method
    std::vector<V3> nbv;
    std::vector<double> bdens, bpress;
    std::vector<V3> pbn;

    size_t N() const { return x.size(); }

    // Appends one particle to the active set.  The reference's arrays are
    // sized for every emitter's budget up front and read only below the
    // active count, so a new particle's fields start at zero (k, kappa_v
    // included) and its position and velocity are the emitter's.
    void appendParticle(V3 pos, V3 vel) {
        x.push_back(pos); v.push_back(vel); a.push_back(V3()); accStep.push_back(V3());
        rho.push_back(0.0); p.push_back(0.0);
        alpha.push_back(0.0); divErr.push_back(0.0); rhoStar.push_back(0.0); k.push_back(0.0); kv.push_back(0.0);
        predR.push_back(V3()); predV.push_back(V3()); pacc.push_back(V3());
        normal.push_back(V3());
        pbn.push_back(V3());
    }
    // Simulation::emitParticles (src/Simulation.cpp:325-329) ->
    // Emitter::emitParticles (src/Emitter.cpp:85-97) and emitSquare /
    // emitCircle (:99-162), which are the same: when time >= nextTimeEmit
    // and a whole group is left in the budget, the group is emitted with
    // velocity (0, 0, v) and position group + t (velocity + g t), t = time -
    // nextTimeEmit, both rotated by the quaternion, then offset by the
    // emitter position; nextTimeEmit advances by spacing 1.105 dist / v.
    void emitParticles() {
        for (auto& em : emitters) {
            if (!(time >= em.nextTimeEmit)) continue;
            const double timeOffset = time - em.nextTimeEmit;
            const double dist = 2.0 * r;
            if (em.group.size() <= em.remaining) {
                for (size_t i = 0; i < em.group.size(); ++i) {
                    V3 velocity(0, 0, em.e.velocity);
                    V3 position = em.group[i] + (velocity + g * timeOffset) * timeOffset;
                    position = quatRotate(em.e.rotation, position);
                    velocity = quatRotate(em.e.rotation, velocity);
                    position += V3(em.e.position[0], em.e.position[1], em.e.position[2]);
                    appendParticle(position, velocity);
                }
                em.remaining -= (unsigned)em.group.size();
                em.nextTimeEmit += em.e.spacing * 1.105 * dist / em.e.velocity;
            }
        }
    }

    // ------------------------------------------------------------ grid
    // src/HashTable.h:176-196 clear(): empties per-cell point and neighbor
    // lists but keeps the cells themselves.
    void clearFluid() {
        for (auto& c : table) {
            c.second.points.clear();
            c.second.neighbors.clear();
            c.second.bneighbors.clear();
        }
    }
    // src/SPHSolver.cpp:178-203, src/HashTable.h:79-86: serial insert in id order.
    void insertFluid() {
        clearFluid();
        for (uint32_t i = 0; i < N(); ++i) table[key_of(x[i], h)].points.push_back(PointInfo{i, 0});
    }
    // src/SPHSolver.cpp:205-221, src/HashTable.h:94-102.
    void insertBoundary() {
        for (auto& c : table) c.second.bpoints.clear();
        for (uint32_t i = 0; i < xb.size(); ++i) table[key_of(xb[i], h)].bpoints.push_back(PointInfo{i, 0});
    }
    // src/HashTable.h:136-169: per cell, concatenate the 27 neighbor cells'
    // lists with x outermost and z innermost, each offset from -1 to 1.
    void neighborhoodSearch() {
        std::vector<Key> keys;
        keys.reserve(table.size());
        for (auto& c : table) keys.push_back(c.first);
        #pragma omp parallel for schedule(dynamic, 64)
        for (long ci = 0; ci < (long)keys.size(); ++ci) {
            Key key = keys[ci];
            Bucket& cell = table.find(key)->second;
            for (int dx = -1; dx <= 1; ++dx)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dz = -1; dz <= 1; ++dz) {
                        auto it = table.find(Key{key.x + dx, key.y + dy, key.z + dz});
                        if (it != table.end()) {
                            cell.neighbors.insert(cell.neighbors.end(), it->second.points.begin(), it->second.points.end());
                            cell.bneighbors.insert(cell.bneighbors.end(), it->second.bpoints.begin(), it->second.bpoints.end());
                        }
                    }
        }
    }
    const Bucket& cellOf(V3 p) const {
        static const Bucket empty;
        auto it = table.find(key_of(p, h));
        return it == table.end() ? empty : it->second;
    }

    // --------------------------------------------------------- densities
    // src/SPHSolver.cpp:44-91 (Akinci branch at :78-89)
    void computeDensities() {
        const long n = (long)N();
        #pragma omp parallel for
        for (long i = 0; i < n; ++i) {
            const V3 ri = x[i];
            const Bucket& c = cellOf(ri);
            double density = 0;
            for (const PointInfo& pj : c.neighbors) density += m * K.W(ri - x[pj.id]);
            for (const PointInfo& pb : c.bneighbors) density += rho0 * vb[pb.id] * K.W(ri - xb[pb.id]);
            rho[i] = density;
        }
    }
    // src/AkinciBoundayModel.cpp:11-40
    void computeVolume() {
        const long nb = (long)xb.size();
        #pragma omp parallel for
        for (long i = 0; i < nb; ++i) {
            const V3 ri = xb[i];
            const Bucket& c = cellOf(ri);
            double delta_ = 0.0;
            for (const PointInfo& pb : c.bneighbors) delta_ += K.W(ri - xb[pb.id]);
            vb[i] = 1.0 / delta_;
        }
    }

    // src/SPHSolver.cpp:93-121: PCISPH boundary particle densities from the
    // fluid candidates of the boundary particle's cell and the boundary ones
    // (boundary mass m_b = rho0 (2r)^3, src/Simulation.cpp:465-467).
    void computeBoundaryDensities() {
        const long nb = (long)xb.size();
        const double mb = rho0 * vbPcisph();
        #pragma omp parallel for
        for (long b = 0; b < nb; ++b) {
            const V3 ri = xb[b];
            const Bucket& c = cellOf(ri);
            double density = 0;
            for (const PointInfo& pj : c.neighbors) density += m * K.W(ri - x[pj.id]);
            for (const PointInfo& pc : c.bneighbors) density += mb * K.W(ri - xb[pc.id]);
            bdens[b] = density;
        }
    }
    // The PCISPH boundary particles' volume-equivalent: with V_b = (2r)^3
    // the Akinci-form terms rho0 V_b (...) are the method's m_b (...) terms,
    // evaluated in the same order (src/DFSPHSolver.cpp:149-160 and the six
    // other DFSPH sites, src/SPHSolver.cpp:72-80, src/PCISPHSolver.cpp:193-202).
    double vbPcisph() const { return std::pow(2.0 * r, 3.0); }

    // glm::normalize: v * (1 / sqrt(dot(v, v))) (NaN for v = 0)
    static V3 normalize(V3 v) { return v * (1.0 / std::sqrt(dot(v, v))); }

    // PCISPHBoundaryModel::correctPositions (src/PCISPHBoundaryModel.cpp:
    // 219-309) and correctPredPositions (:385-439): for every boundary
    // candidate of the cell of the current position r_i with |r_i - r_b| <
    // h/2, w = max(0, (h/2 - |q - r_b|) / (h/2)) with q = r_i (WCSPH, DFSPH)
    // or the predicted position (PCISPH); n_i = normalize(sum w n_b) is
    // stored (NaN when the sum is zero) and, when sum w > 0, the target
    // (r_i, or the predicted position for `pred`) moves by
    // (1 / sum w) (sum w (h/2 - |q - r_b|)) n_i.
    void pbCorrect(bool usePred, bool movePred) {
        const long n = (long)N();
        const double restRadius = 0.5 * h;
        #pragma omp parallel for
        for (long i = 0; i < n; ++i) {
            const V3 ri = x[i];
            const Bucket& c = cellOf(ri);
            V3 avg;
            double wsum = 0, w2sum = 0;
            for (const PointInfo& pb_ : c.bneighbors) {
                const V3 rb = xb[pb_.id];
                const V3 q = (usePred ? predR[i] : ri) - rb;
                if (length(ri - rb) < restRadius) {
                    const double lq = length(q);
                    const double w = std::max(0.0, (restRadius - lq) / restRadius);
                    wsum += w;
                    w2sum += w * (restRadius - lq);
                    avg += w * nbv[pb_.id];
                }
            }
            const V3 ni = normalize(avg);
            pbn[i] = ni;
            if (wsum > 0) {
                const V3 dlt = ((1.0 / wsum) * w2sum) * ni;
                if (movePred) predR[i] += dlt;
                else x[i] += dlt;
            }
        }
    }
    // PCISPHBoundaryModel::correctVelocities (src/PCISPHBoundaryModel.cpp:
    // 311-383): where |n_i| > 0, v_i = tangentialFct (u - (u.n) n) -
    // normalFct (u.n) n with u = v_i (WCSPH, DFSPH) or the predicted
    // velocity (PCISPH).
    void pbCorrectVelocities(bool usePred) {
        const long n = (long)N();
        const double nf = d.pb_normal_fct, tf = d.pb_tangential_fct;
        #pragma omp parallel for
        for (long i = 0; i < n; ++i) {
            const V3 ni = pbn[i];
            if (length(ni) > 0) {
                const V3 u = usePred ? predV[i] : v[i];
                const V3 nVel = dot(u, ni) * ni;
                const V3 tVel = u - nVel;
                v[i] = tf * tVel - nf * nVel;
            }
        }
    }

    // ---------------------------------------------------- NP forces
    // src/XSPHViscosity.h:32-69
    void xsph() {
        const long n = (long)N();
        const double nu = d.viscosity, nub = d.boundary_viscosity;
        const double dt = ts;
        #pragma omp parallel for
        for (long i = 0; i < n; ++i) {
            const V3 ri = x[i], vi = v[i];
            const Bucket& c = cellOf(ri);
            V3 ai = a[i];
            for (const PointInfo& pj : c.neighbors) {
                const uint32_t j = pj.id;
                const double rj = planted(PL_XSPH_RHO_I) ? rho[i] : rho[j];
                if (planted(PL_XSPH_NO_DT)) ai += nu * (m / rj) * (v[j] - vi) * K.W(ri - x[j]);
                else ai += nu * (m / rj) * (v[j] - vi) * K.W(ri - x[j]) / dt;
            }
            if (nub != 0.0 && !pbm) {
                for (const PointInfo& pb : c.bneighbors) {
                    const uint32_t b = pb.id;
                    ai += nub * (rho0 * vb[b] / rho[i]) * (-vi) * K.W(ri - xb[b]) / dt;
                }
            }
            a[i] = ai;
        }
    }
    // ViscoK::laplW (src/ViscoK.h:28-35): 45/(pi h^6) (h - |r|) for |r| <= h
    double laplW(const V3& r) const {
        const double mr = length(r);
        return (mr <= h) ? (45.0 / (M_PI * h * h * h * h * h * h)) * (h - mr) : 0.0;
    }
    // src/StandardViscosityForce.h:30-76
    void standardViscosity() {
        const long n = (long)N();
        const double nu = d.viscosity, nub = d.boundary_viscosity;
        #pragma omp parallel for
        for (long i = 0; i < n; ++i) {
            const V3 ri = x[i], vi = v[i];
            const Bucket& c = cellOf(ri);
            V3 f;
            for (const PointInfo& pj : c.neighbors) {
                const uint32_t j = pj.id;
                f += m * (v[j] - vi) / rho[j] * laplW(ri - x[j]);
            }
            f *= nu / rho0;
            V3 ai = a[i] + f;
            if (nub != 0.0 && !pbm) {
                V3 fb;
                for (const PointInfo& pb : c.bneighbors) {
                    const uint32_t b = pb.id;
                    fb += rho0 * vb[b] * (-1.0 * vi) / rho[i] * laplW(ri - xb[b]);
                }
                fb *= nub / rho0;
                ai += fb;
            }
            a[i] = ai;
        }
    }
    // src/ArtificialViscosity.h:30-84
    void artificialViscosity() {
        const long n = (long)N();
        const double nu = d.viscosity, nub = d.boundary_viscosity;
        const double h2 = h * h;
        #pragma omp parallel for
        for (long i = 0; i < n; ++i) {
            const V3 ri = x[i], vi = v[i];
            const double dens_i = rho[i];
            const Bucket& c = cellOf(ri);
            V3 ai = a[i];
            for (const PointInfo& pj : c.neighbors) {
                const uint32_t j = pj.id;
                const V3 rirj = ri - x[j];
                const double dotProd = dot(vi - v[j], rirj);
                const double dens_j = rho[j];
                const double sqDist = length(rirj) * length(rirj);
                const double kij = 2 * rho0 / (dens_i + dens_j);
                ai += kij * 10.0 * 0.1 * nu * (m / dens_j) * dotProd / (sqDist + 0.01 * h2) * K.gradW(rirj);
            }
            if (nub != 0.0 && !pbm) {
                for (const PointInfo& pb : c.bneighbors) {
                    const uint32_t b = pb.id;
                    const V3 rirb = ri - xb[b];
                    const double dotProd = dot(vi, rirb);
                    const double sqDist = length(rirb) * length(rirb);
                    ai += 10.0 * 0.1 * nub * (rho0 * vb[b] / dens_i) * dotProd / (sqDist + 0.01 * h2) * K.gradW(rirb);
                }
            }
            a[i] = ai;
        }
    }
    // src/SurfaceTensionAkinci.h:34-98
    void surfaceTension() {
        const long n = (long)N();
        const double st = d.surface_tension;
        #pragma omp parallel for
        for (long i = 0; i < n; ++i) {
            const V3 ri = x[i];
            const Bucket& c = cellOf(ri);
            V3 ni;
            for (const PointInfo& pj : c.neighbors) {
                const uint32_t j = pj.id;
                double vol = m / rho[j];
                ni += vol * K.gradW(ri - x[j]);
            }
            ni *= 1.0 * h;
            normal[i] = ni;
        }
        #pragma omp parallel for
        for (long i = 0; i < n; ++i) {
            const V3 ri = x[i];
            const V3 ni = normal[i];
            const double dens_i = rho[i];
            const Bucket& c = cellOf(ri);
            V3 ai = a[i];
            for (const PointInfo& pj : c.neighbors) {
                const uint32_t j = pj.id;
                const V3 rij = ri - x[j];
                const double rijMag = length(rij);
                const double kij = (planted(PL_ST_K_HALF) ? 1.0 : 2.0) * rho0 / (dens_i + rho[j]);
                if (rijMag > 0.0 && rijMag <= h) {
                    V3 cohesion = -m * K.cohesion(rij) * rij / rijMag;
                    V3 curvature = -1.0 * (planted(PL_ST_NPLUS) ? ni + normal[j] : ni - normal[j]);
                    ai += st * kij * (cohesion + curvature);
                }
            }
            a[i] = ai;
        }
    }
    // src/AdhesionForce.h:30-62
    void adhesionForce() {
        const long n = (long)N();
        const double beta = d.adhesion;
        #pragma omp parallel for
        for (long i = 0; i < n; ++i) {
            const V3 ri = x[i];
            const Bucket& c = cellOf(ri);
            V3 ai = a[i];
            for (const PointInfo& pb : c.bneighbors) {
                const uint32_t b = pb.id;
                V3 rib = ri - xb[b];
                double mod_rib = length(rib);
                const double c = (planted(PL_ADH_COEF) ? K.adc : 1.0) * (planted(PL_ADH_NO_RHO0) ? 1.0 : rho0);
                if (mod_rib > 0.0) ai -= beta * c * vb[b] * K.adhesion(rib) * rib / mod_rib;
            }
            a[i] = ai;
        }
    }
    // Creation order of the force objects is viscosity, surface tension,
    // adhesion (src/Simulation.cpp:414-423), and Simulation::
    // computeNonPressureForces runs them in that order (:313-323).
    void computeNonPressureForces() {
        if (d.viscosity_method == 0 && d.viscosity != 0.0) standardViscosity();
        if (d.viscosity_method == 1 && d.viscosity != 0.0) artificialViscosity();
        if (d.viscosity_method == 2 && d.viscosity != 0.0) xsph();
        if (d.surface_tension_method == 3 && d.surface_tension != 0.0) surfaceTension();
        if (d.adhesion_method == 4 && d.adhesion != 0.0) adhesionForce();
    }

    // ------------------------------------------------------------ WCSPH
    // src/WCSPHSolver.cpp:80-95
    void wcsphPressures() {
        const long n = (long)N();
        const double kappa = d.stiffness, gam = d.gamma;
        #pragma omp parallel for
        for (long i = 0; i < n; ++i)
            p[i] = std::max(0.0, kappa * rho0 / gam * (std::pow(rho[i] / rho0, gam) - 1));
        // boundary pressures of the PCISPH method (src/WCSPHSolver.cpp:97-114)
        if (pbm)
            for (size_t b = 0; b < xb.size(); ++b)
                bpress[b] = std::max(0.0, kappa * rho0 / gam * (std::pow(bdens[b] / rho0, gam) - 1));
    }
    // src/WCSPHSolver.cpp:132-188 (Akinci branch :174-186)
    void wcsphPressureForce() {
        const long n = (long)N();
        #pragma omp parallel for
        for (long i = 0; i < n; ++i) {
            const V3 ri = x[i];
            const Bucket& c = cellOf(ri);
            V3 ai = a[i];
            const double densPress_i = planted(PL_WC_P_OVER_RHO) ? p[i] / rho[i] : p[i] / (rho[i] * rho[i]);
            for (const PointInfo& pj : c.neighbors) {
                const uint32_t j = pj.id;
                double dpj = planted(PL_WC_P_OVER_RHO) ? p[j] / rho[j] : p[j] / (rho[j] * rho[j]);
                if (planted(PL_WC_PJ_AS_PI)) dpj = densPress_i;
                ai -= m * (densPress_i + dpj) * K.gradW(ri - x[j]);
            }
            if (pbm) {   // src/WCSPHSolver.cpp:160-172
                const double mb = rho0 * vbPcisph();
                for (const PointInfo& pc : c.bneighbors) {
                    const uint32_t b = pc.id;
                    ai -= mb * (densPress_i + bpress[b] / (bdens[b] * bdens[b])) * K.gradW(ri - xb[b]);
                }
            } else {
                for (const PointInfo& pc : c.bneighbors) {
                    const uint32_t b = pc.id;
                    ai -= rho0 * vb[b] * (densPress_i * (planted(PL_WC_BOUNDARY_SYM) ? 2.0 : 1.0)) * K.gradW(ri - xb[b]);
                }
            }
            a[i] = ai;
        }
    }
    // src/SPHSolver.cpp:141-170 (maxVel kept as a proper reduction)
    void integrate() {
        const long n = (long)N();
        const double dt = ts;
        double mv = 0.0;
        #pragma omp parallel for reduction(max : mv)
        for (long i = 0; i < n; ++i) {
            accStep[i] = a[i];
            V3 ai = a[i] + g;
            v[i] += ai * dt;
            x[i] += v[i] * dt;
            a[i] = V3();
            mv = std::max(mv, length(v[i]));
        }
        maxVel = mv;
    }
    // CubeBoundaryModel::correctPositionAndVelocity (src/CubeBoundaryModel.cpp:16-94)
    void cubeCorrect() {
        if (!d.cube_enabled) return;
        const V3 mn(d.cube_min[0], d.cube_min[1], d.cube_min[2]);
        const V3 mx(d.cube_max[0], d.cube_max[1], d.cube_max[2]);
        const double normalFct = d.cube_normal_fct, tangentialFct = d.cube_tangential_fct;
        const long n = (long)N();
        #pragma omp parallel for
        for (long i = 0; i < n; ++i) {
            V3& pos = x[i];
            V3& vel = v[i];
            if (pos.x < mn.x) {
                pos.x = mn.x;
                if (vel.x < 0.f) vel.x *= -normalFct;
                vel.y *= tangentialFct;
                vel.z *= tangentialFct;
            } else if (pos.x > mx.x) {
                pos.x = mx.x;
                if (vel.x > 0.f) vel.x *= -normalFct;
                vel.y *= tangentialFct;
                vel.z *= tangentialFct;
            }
            if (pos.y < mn.y) {
                pos.y = mn.y;
                if (vel.y < 0.f) vel.y *= -normalFct;
                vel.x *= tangentialFct;
                vel.z *= tangentialFct;
            } else if (pos.y > mx.y) {
                pos.y = mx.y;
                if (vel.y > 0.f) vel.y *= -normalFct;
                vel.x *= tangentialFct;
                vel.z *= tangentialFct;
            }
            if (pos.z < mn.z) {
                pos.z = mn.z;
                if (vel.z < 0.f) vel.z *= -normalFct;
                vel.x *= tangentialFct;
                vel.y *= tangentialFct;
            } else if (pos.z > mx.z) {
                pos.z = mx.z;
                if (vel.z > 0.f) vel.z *= -normalFct;
                vel.x *= tangentialFct;
                vel.y *= tangentialFct;
            }
        }
    }

    // src/WCSPHSolver.cpp:14-78
    void wcsphStep() {
        ++steps;
        insertFluid();
        neighborhoodSearch();
        computeDensities();
        if (pbm) computeBoundaryDensities();   // src/SPHSolver.cpp:123-135
        wcsphPressures();
        wcsphPressureForce();
        computeNonPressureForces();
        integrate();
        cubeCorrect();   // :51-60
        if (pbm) {       // :62-73
            pbCorrect(false, false);
            pbCorrectVelocities(false);
        }
        emitParticles(); // :75
        time += ts;
    }

    // ------------------------------------------------------------ PCISPH
    // src/PCISPHSolver.cpp:389-421
    void computeScalingFactor() {
        const double dist = 2.0 * r;
        V3 sumGrad;
        double sumSqGrad = 0;
        for (int i = -1; i <= 1; i++)
            for (int j = -1; j <= 1; j++)
                for (int kk = -1; kk <= 1; kk++) {
                    V3 rij((double)-i, (double)-j, (double)-kk);
                    rij *= dist;
                    V3 grad = K.poly6Grad(rij);
                    sumGrad += grad;
                    sumSqGrad += dot(grad, grad);
                }
        double beta = 2 * std::pow(m * ts / rho0, 2);
        double dotSumGrad = dot(sumGrad, sumGrad);
        delta = -1.0 / (beta * (-dotSumGrad - sumSqGrad));
    }
    // src/PCISPHSolver.cpp:122-149
    void predictVelocityAndPosition() {
        const long n = (long)N();
        const double dt = ts;
        #pragma omp parallel for
        for (long i = 0; i < n; ++i) {
            V3 acc = a[i] + pacc[i] + g;
            predV[i] = v[i] + acc * dt;
            predR[i] = x[i] + predV[i] * dt;
        }
    }
    // src/PCISPHSolver.cpp:164-212: predicted positions, candidate cells from
    // the current positions (:181).
    void predictFluidDensities() {
        const long n = (long)N();
        #pragma omp parallel for
        for (long i = 0; i < n; ++i) {
            const V3 ri = predR[i];
            const Bucket& c = cellOf(x[i]);
            double density = 0;
            for (const PointInfo& pj : c.neighbors) density += m * K.W(ri - predR[pj.id]);
            for (const PointInfo& pb : c.bneighbors) density += rho0 * vb[pb.id] * K.W(ri - xb[pb.id]);
            rho[i] = density;
        }
    }
    // src/PCISPHSolver.cpp:214-247: PCISPH boundary densities from the
    // predicted fluid positions, with the Poly6 kernel.
    void predictBoundaryDensities() {
        const long nb = (long)xb.size();
        const double mb = rho0 * vbPcisph();
        #pragma omp parallel for
        for (long b = 0; b < nb; ++b) {
            const V3 ri = xb[b];
            const Bucket& c = cellOf(ri);
            double density = 0;
            for (const PointInfo& pj : c.neighbors) density += m * K.poly6W(ri - predR[pj.id]);
            for (const PointInfo& pc : c.bneighbors) density += mb * K.poly6W(ri - xb[pc.id]);
            bdens[b] = density;
        }
    }
    // src/PCISPHSolver.cpp:249-295, with the max taken as a reduction;
    // boundary pressures of the PCISPH method (:276-293, fluid 0's delta).
    void updatePressure() {
        const long n = (long)N();
        double me = 0;
        #pragma omp parallel for reduction(max : me)
        for (long i = 0; i < n; ++i) {
            double error = planted(PL_PCI_NO_CLAMP) ? rho[i] - rho0 : std::max(0.0, rho[i] - rho0);
            p[i] += delta * error;
            me = std::max(me, error / rho0);
        }
        maxDensityError = me;
        if (pbm)
            for (size_t b = 0; b < xb.size(); ++b) bpress[b] += delta * std::max(0.0, bdens[b] - rho0);
    }
    // src/PCISPHSolver.cpp:297-358 (Akinci branch :342-353)
    void computePressureAcc() {
        const long n = (long)N();
        #pragma omp parallel for
        for (long i = 0; i < n; ++i) {
            const V3 ri = x[i];
            const Bucket& c = cellOf(ri);
            V3 pa;
            const bool pr = planted(PL_PCI_P_OVER_RHO);
            auto grad = [&](V3 rr) { return planted(PL_PCI_CUBIC) ? K.gradW(rr) : K.spikyGrad(rr); };
            const double densPress_i = pr ? p[i] / rho[i] : p[i] / (rho[i] * rho[i]);
            for (const PointInfo& pj : c.neighbors) {
                const uint32_t j = pj.id;
                pa -= m * (densPress_i + (pr ? p[j] / rho[j] : p[j] / (rho[j] * rho[j]))) * grad(ri - x[j]);
            }
            if (pbm) {   // src/PCISPHSolver.cpp:330-340
                const double mb = rho0 * vbPcisph();
                for (const PointInfo& pc : c.bneighbors) {
                    const uint32_t b = pc.id;
                    pa -= mb * (densPress_i + bpress[b] / (bdens[b] * bdens[b])) * K.spikyGrad(ri - xb[b]);
                }
            } else {
                for (const PointInfo& pc : c.bneighbors) {
                    const uint32_t b = pc.id;
                    pa -= rho0 * vb[b] * (densPress_i) * grad(ri - xb[b]);
                }
            }
            pacc[i] = pa;
        }
    }
    // src/PCISPHSolver.cpp:360-387
    void pressureSolver() {
        iterations = 0;
        while ((maxDensityError > d.eta || iterations < d.min_iterations) && iterations < d.max_iterations) {
            predictVelocityAndPosition();
            if (pbm) pbCorrect(true, true);        // correctPredPositions (:370-377)
            predictFluidDensities();
            if (pbm) predictBoundaryDensities();   // predictDensities (:151-162)
            updatePressure();
            computePressureAcc();
            ++iterations;
        }
        sumIt += iterations;
    }
    // src/PCISPHSolver.cpp:444-474
    void pcisphIntegrate() {
        const long n = (long)N();
        const double dt = ts;
        double mv = 0.0;
        #pragma omp parallel for reduction(max : mv)
        for (long i = 0; i < n; ++i) {
            accStep[i] = a[i];
            V3 ai = a[i] + (g + pacc[i]);  // a += g + pacc (:465)
            v[i] += ai * dt;
            x[i] += v[i] * dt;
            a[i] = V3();
            mv = std::max(mv, length(v[i]));
        }
        maxVel = mv;
    }
    // src/PCISPHSolver.cpp:21-81, initPressure :83-120
    void pcisphStep() {
        ++steps;
        insertFluid();
        neighborhoodSearch();
        computeDensities();
        if (pbm) computeBoundaryDensities();
        computeNonPressureForces();
        for (size_t i = 0; i < N(); ++i) { p[i] = 0.0; pacc[i] = V3(); }
        if (pbm) std::fill(bpress.begin(), bpress.end(), 0.0);   // initPressure (:105-119)
        pressureSolver();
        pcisphIntegrate();
        cubeCorrect();   // src/PCISPHSolver.cpp:54-63
        if (pbm) {       // :65-76: with the loop's last predicted position and velocity
            pbCorrect(true, false);
            pbCorrectVelocities(true);
        }
        emitParticles(); // :78
        time += ts;
    }

    // ------------------------------------------------------------- DFSPH
    // src/DFSPHSolver.cpp:114-189 (Akinci branch :161-176)
    void computeAlpha() {
        const long n = (long)N();
        #pragma omp parallel for
        for (long i = 0; i < n; ++i) {
            const V3 ri = x[i];
            const Bucket& c = cellOf(ri);
            V3 sumGrad;
            double sumSqGrad = 0.0;
            for (const PointInfo& pj : c.neighbors) {
                V3 massGrad = m * K.gradW(ri - x[pj.id]);
                double mag = length(massGrad);
                sumGrad += massGrad;
                sumSqGrad += mag * mag;
            }
            for (const PointInfo& pb : c.bneighbors) {
                V3 massGrad = rho0 * vb[pb.id] * K.gradW(ri - xb[pb.id]);
                sumGrad += massGrad;
            }
            double sumGradMag = length(sumGrad);
            double denominator = sumGradMag * sumGradMag + (planted(PL_ALPHA_NO_SUMSQ) ? 0.0 : sumSqGrad);
            alpha[i] = (std::fabs(denominator) > eps) ? rho[i] / denominator : 0.0;
        }
    }
    // src/DFSPHSolver.cpp:191-214
    void predictVelocities() {
        const long n = (long)N();
        const double dt = ts;
        #pragma omp parallel for
        for (long i = 0; i < n; ++i) {
            accStep[i] = a[i];
            V3 acc = a[i] + g;
            v[i] += acc * dt;
            a[i] = V3();
        }
    }
    // src/DFSPHSolver.cpp:216-236
    void updatePositions() {
        const long n = (long)N();
        const double dt = ts;
        #pragma omp parallel for
        for (long i = 0; i < n; ++i) {
            x[i] += v[i] * dt;
            if (f32Positions) x[i] = V3((float)x[i].x, (float)x[i].y, (float)x[i].z);
        }
    }
    // src/DFSPHSolver.cpp:238-309 (Akinci branch :285-295)
    void predictDensities() {
        const long n = (long)N();
        const double dt = ts;
        double sum = 0.0;
        #pragma omp parallel for reduction(+ : sum)
        for (long i = 0; i < n; ++i) {
            const V3 ri = x[i], vi = v[i];
            const Bucket& c = cellOf(ri);
            double densityChange = 0;
            for (const PointInfo& pj : c.neighbors) {
                const uint32_t j = pj.id;
                densityChange += m * dot(vi - v[j], K.gradW(ri - x[j]));
            }
            for (const PointInfo& pb : c.bneighbors) {
                const uint32_t b = pb.id;
                densityChange += rho0 * vb[b] * dot(vi, K.gradW(ri - xb[b]));
            }
            double pd = rho[i] + densityChange * dt;
            pd = std::max(pd, rho0);
            rhoStar[i] = pd;
            sum += (pd - rho0) / rho0;
        }
        avgDensError = (n == 0) ? 0.0 : sum / (double)n;
    }
    // shared velocity sweep of correctDensityError (:335-395) and
    // correctDivergenceError (:520-579): v_i -= dt * sum.
    void velocitySweep(const std::vector<double>& err, double scale, std::vector<double>& acc) {
        const long n = (long)N();
        const double dt = ts;
        #pragma omp parallel for
        for (long i = 0; i < n; ++i) {
            const V3 ri = x[i];
            const double kx = planted(PL_KAPPA_X2) ? 2.0 : 1.0;
            const bool r0 = planted(PL_KAPPA_RHO0);
            const double ki = kx * err[i] * alpha[i] / (scale * (r0 ? rho0 : rho[i]));
            acc[i] += planted(PL_K_NO_ALPHA) ? err[i] : err[i] * alpha[i];
            const Bucket& c = cellOf(ri);
            V3 sum;
            for (const PointInfo& pj : c.neighbors) {
                const uint32_t j = pj.id;
                const double kj = kx * err[j] * alpha[j] / (scale * (r0 ? rho0 : rho[j]));
                const double kSum = ki + kj;
                if (std::fabs(kSum) > eps) sum += m * kSum * K.gradW(ri - x[j]);
            }
            if (std::fabs(ki) > eps) {
                for (const PointInfo& pb : c.bneighbors) {
                    const uint32_t b = pb.id;
                    sum += rho0 * vb[b] * ki * K.gradW(ri - xb[b]);
                }
            }
            if (planted(PL_SWEEP_SIGN)) v[i] += dt * sum;
            else v[i] -= dt * sum;
        }
    }
    // src/DFSPHSolver.cpp:311-406.  The density error passed to the sweep is
    // (rho* - rho0); its kappa is that times alpha / (dt^2 rho) (:346).
    void correctDensityError() {
        const double dt = ts;
        densityWarmStart();
        predictDensities();
        avgDensError0 = avgDensError;
        std::vector<double> err(N());
        iterations = 0;
        avgDensError = 0;
        while (((avgDensError > d.eta) || (iterations < d.min_iterations)) && iterations < d.max_iterations) {
            for (size_t i = 0; i < N(); ++i) err[i] = rhoStar[i] - rho0;
            velocitySweep(err, planted(PL_KAPPA_DT) ? dt : dt * dt, k);
            predictDensities();
            ++iterations;
        }
        sumIt += iterations;
    }
    // src/DFSPHSolver.cpp:408-492 (Akinci branch :466-478, cutoff :480-483)
    void computeDivergenceError() {
        const long n = (long)N();
        double sum = 0.0;
        #pragma omp parallel for reduction(+ : sum)
        for (long i = 0; i < n; ++i) {
            const V3 ri = x[i], vi = v[i];
            const Bucket& c = cellOf(ri);
            double de = 0;
            unsigned numNeighbors = 0;
            for (const PointInfo& pj : c.neighbors) {
                const uint32_t j = pj.id;
                V3 grad = K.gradW(ri - x[j]);
                de += m * dot(vi - v[j], grad);
                if (length(grad) > 0.0) numNeighbors++;
            }
            for (const PointInfo& pb : c.bneighbors) {
                const uint32_t b = pb.id;
                V3 grad = K.gradW(ri - xb[b]);
                de += rho0 * vb[b] * dot(vi, grad);
            }
            if (numNeighbors < 20) de = 0.0;
            else de = std::max(de, 0.0);
            divErr[i] = de;
            sum += de / rho0;
        }
        avgDivError = sum / (double)n;
    }
    // src/DFSPHSolver.cpp:494-590; the loop bound is maxIterations (:512).
    void correctDivergenceError() {
        const double dt = ts;
        computeDivergenceError();
        avgDivError0 = avgDivError;
        iterationsV = 0;
        avgDivError = 0;
        const unsigned bound = planted(PL_DIV_LOOP_MAXV) ? d.max_iterations_v : d.max_iterations;
        while (((avgDivError > d.eta_v / dt) || (iterationsV < d.min_iterations_v)) && iterationsV < bound) {
            velocitySweep(divErr, planted(PL_KV_DT2) ? dt * dt : dt, kv);
            computeDivergenceError();
            ++iterationsV;
        }
        sumItV += iterationsV;
    }
    // src/DFSPHSolver.cpp:592-690 (Akinci branch :661-674 uses 2 k_i)
    void densityWarmStart() {
        const long n = (long)N();
        const double dt = ts, dt2 = dt * dt;
        #pragma omp parallel for
        for (long i = 0; i < n; ++i) k[i] = -std::max(-k[i] / rho[i], -0.000000001) / dt2;
        #pragma omp parallel for
        for (long i = 0; i < n; ++i) {
            const V3 ri = x[i];
            const double ki = k[i];
            const Bucket& c = cellOf(ri);
            V3 sum;
            for (const PointInfo& pj : c.neighbors) {
                const uint32_t j = pj.id;
                const double kSum = ki + k[j];
                if (std::fabs(kSum) > eps) sum += m * kSum * K.gradW(ri - x[j]);
            }
            for (const PointInfo& pb : c.bneighbors) {
                const uint32_t b = pb.id;
                const double kSum = 2.0 * ki;
                if (std::fabs(kSum) > eps) sum += rho0 * vb[b] * kSum * K.gradW(ri - xb[b]);
            }
            v[i] -= sum * dt;
        }
        #pragma omp parallel for
        for (long i = 0; i < n; ++i) k[i] = 0.0;
    }
    // src/DFSPHSolver.cpp:827-870 (sqrt of the speed, floor 0.1)
    void updateTimeStep() {
        const long n = (long)N();
        const double dt = ts;
        double mv = 0.1;
        #pragma omp parallel for reduction(max : mv)
        for (long i = 0; i < n; ++i) {
            V3 acc = planted(PL_DT_NO_G) ? a[i] : a[i] + g;
            V3 cur = v[i] + acc * dt;
            mv = std::max(mv, planted(PL_DT_ABS_V) ? length(cur) : std::sqrt(length(cur)));
        }
        maxVel = mv;
        double newTs;
        double cflTs = d.cfl_factor * 0.4 * 2.0 * r / mv;
        lastCfl = cflTs;
        if (cflTs < dt) newTs = cflTs;
        else if (iterations > 3 || (planted(PL_DT_099_GE3) && iterations == 3)) newTs = dt * 0.99;
        else newTs = dt * 1.001;
        newTs = std::max(newTs, minTs);
        newTs = std::min(newTs, maxTs);
        ts = newTs;
    }
    // src/DFSPHSolver.cpp:38-112
    void dfsphStep() {
        steps++;
        computeNonPressureForces();
        updateTimeStep();
        predictVelocities();
        if (pbm) pbCorrectVelocities(false);   // src/DFSPHSolver.cpp:51-58 (normals of the last step)
        correctDensityError();
        updatePositions();
        if (pbm) pbCorrect(false, false);      // :65-75
        emitParticles();   // src/DFSPHSolver.cpp:77
        insertFluid();
        neighborhoodSearch();
        computeDensities();
        computeAlpha();
        correctDivergenceError();
        cubeCorrect();   // src/DFSPHSolver.cpp:99-108
        time += ts;
    }

    // ------------------------------------------------------------ set-up
    void setDesc(const oracle_desc& dd) {
        d = dd;
        pbm = d.boundary_method == 1;
        r = d.particle_radius;
        h = 4.0 * r;                      // src/Simulation.cpp:197
        rho0 = d.density0;
        m = rho0 * std::pow(2.0 * r, 3.0);  // src/Simulation.cpp:134-146
        g = V3(d.gravity[0], d.gravity[1], d.gravity[2]);
        ts = d.time_step;
        fps = d.fps;
        double minFps = 1.0 / fps;        // src/TimeManager.h:55
        minTs = minFps < d.min_time_step ? minFps : d.min_time_step;
        maxTs = d.max_time_step;
        K.set(h);
    }
    void resize(size_t n) {
        x.resize(n); v.resize(n); a.assign(n, V3()); accStep.assign(n, V3());
        rho.assign(n, 0.0); p.assign(n, 0.0);
        alpha.assign(n, 0.0); divErr.assign(n, 0.0); rhoStar.assign(n, 0.0); k.assign(n, 0.0); kv.assign(n, 0.0);
        predR.assign(n, V3()); predV.assign(n, V3()); pacc.assign(n, V3());
        normal.assign(n, V3());
        pbn.assign(n, V3());
    }
    // Simulation::init (src/Simulation.cpp:102-110) -> solver->init():
    // WCSPH src/WCSPHSolver.cpp:9-12, PCISPH src/PCISPHSolver.cpp:12-19,
    // DFSPH src/DFSPHSolver.cpp:24-36, base src/SPHSolver.cpp:22-39.
    void init() {
        table.clear();
        table.reserve(2 * N());
        if (d.method == 1) computeScalingFactor();
        insertBoundary();
        neighborhoodSearch();
        if (pbm) {   // no volume pass for the PCISPH method (src/SPHSolver.cpp:22-39)
            vb.assign(xb.size(), vbPcisph());
            bdens.assign(xb.size(), 0.0);
            bpress.assign(xb.size(), 0.0);
            if (nbv.size() != xb.size()) nbv.assign(xb.size(), V3());
        } else {
            computeVolume();
        }
        if (d.method == 2) {
            insertFluid();
            neighborhoodSearch();
            computeDensities();
            computeAlpha();
        }
        initialised = true;
    }
    // Recompute the position-derived state after an injected state: the
    // grid, densities and (DFSPH) alpha, as DFSPHSolver::step leaves them
    // after its divergence solve's grid rebuild (:87-97).
    void refresh() {
        insertFluid();
        neighborhoodSearch();
        computeDensities();
        if (d.method == 2) computeAlpha();
    }
    void step() {
        if (d.method == 0) wcsphStep();
        else if (d.method == 1) pcisphStep();
        else dfsphStep();
    }
};

void put3(const std::vector<V3>& src, double* out) {
    for (size_t i = 0; i < src.size(); ++i) { out[3 * i] = src[i].x; out[3 * i + 1] = src[i].y; out[3 * i + 2] = src[i].z; }
}

}  // namespace

extern "C" {

void* oracle_create(const oracle_desc* d) {
    Oracle* o = new Oracle();
    o->setDesc(*d);
    return o;
}
void oracle_destroy(void* h) { delete (Oracle*)h; }

int oracle_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

// Fluid state from FP32 (promoted exactly) or FP64 arrays of n*3 values.
int oracle_set_fluid(void* h, const float* xyz, const float* vel, unsigned n) {
    Oracle* o = (Oracle*)h;
    o->resize(n);
    for (unsigned i = 0; i < n; ++i) {
        o->x[i] = V3(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
        o->v[i] = vel ? V3(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]) : V3();
    }
    return 0;
}
int oracle_set_boundary(void* h, const float* xyz, unsigned nb) {
    Oracle* o = (Oracle*)h;
    o->xb.resize(nb);
    o->vb.assign(nb, 0.0);
    for (unsigned i = 0; i < nb; ++i) o->xb[i] = V3(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    return 0;
}
// PCISPH boundary normals, one per boundary particle (before init).
int oracle_set_boundary_normals(void* h, const float* n, unsigned nb) {
    Oracle* o = (Oracle*)h;
    if (nb != o->xb.size()) return -1;
    o->nbv.resize(nb);
    for (unsigned i = 0; i < nb; ++i) o->nbv[i] = V3(n[3 * i], n[3 * i + 1], n[3 * i + 2]);
    return 0;
}
// The fluid's normal field (FluidModel::n) of an injected state: DFSPH's
// velocity correction at the start of a step reads the last step's.
int oracle_set_pb_normals(void* h, const float* n) {
    Oracle* o = (Oracle*)h;
    for (size_t i = 0; i < o->N(); ++i) o->pbn[i] = V3(n[3 * i], n[3 * i + 1], n[3 * i + 2]);
    return 0;
}

// FluidModel::addEmitter (src/FluidModel.cpp:124-143), before init.
int oracle_add_emitter(void* h, const oracle_emitter* e) {
    Oracle* o = (Oracle*)h;
    Emitter em;
    em.setup(*e, o->r);
    o->emitters.push_back(em);
    return (int)em.group.size();
}
int oracle_init(void* h) { ((Oracle*)h)->init(); return 0; }

// Switch on planted mistake `id` (0 = none) for every oracle of the
// process; returns the number of plant ids, or -1 for an unknown one.
int oracle_plant(int id) {
    if (id < 0 || id >= PL_COUNT) return -1;
    g_plant = id;
    return PL_COUNT;
}

// Inject a state (FP32 arrays, promoted) for one-step parity from a GPU
// state: positions, velocities, warm-start accumulators (may be null), the
// time step and the last density-solver iteration count (used by the DFSPH
// time-step rule, src/DFSPHSolver.cpp:857-860).  Boundary volumes and the
// boundary grid are kept from oracle_init.
int oracle_set_state(void* h, const float* xyz, const float* vel, const float* k, const float* kv,
                     double dt, unsigned last_iterations) {
    Oracle* o = (Oracle*)h;
    const size_t n = o->N();
    for (size_t i = 0; i < n; ++i) {
        o->x[i] = V3(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
        o->v[i] = V3(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
        o->a[i] = V3();
        o->k[i] = k ? (double)k[i] : 0.0;
        o->kv[i] = kv ? (double)kv[i] : 0.0;
    }
    o->ts = dt;
    o->iterations = last_iterations;
    o->refresh();
    return 0;
}

// Rounds the carried particle state (x, v, kappa, kappa_v) to FP32, as a
// solver that stores them in FP32 between steps holds them; the neighbor
// lists, densities and alpha factors are left as they are (unlike
// oracle_set_state, which recomputes them from the given positions).  With
// the cube boundary on, a coordinate inside the box whose nearest FP32 value
// lies outside it takes the nearest FP32 value inside it instead.
int oracle_round_state_f32(void* h) {
    Oracle* o = (Oracle*)h;
    auto r = [](double& a) { a = (double)(float)a; };
    const bool cube = o->d.cube_enabled != 0;
    auto rx = [&](double& a, int c) {
        const double lo = o->d.cube_min[c], hi = o->d.cube_max[c];
        float f = (float)a;
        if (cube && a >= lo && (double)f < lo) f = std::nextafter(f, INFINITY);
        if (cube && a <= hi && (double)f > hi) f = std::nextafter(f, -INFINITY);
        a = (double)f;
    };
    for (size_t i = 0; i < o->N(); ++i) {
        rx(o->x[i].x, 0); rx(o->x[i].y, 1); rx(o->x[i].z, 2);
        r(o->v[i].x); r(o->v[i].y); r(o->v[i].z);
        r(o->k[i]); r(o->kv[i]);
    }
    return 0;
}

// When on, DFSPH's position update (src/DFSPHSolver.cpp:60) rounds the new
// positions to FP32, so the neighbor search, densities and alpha factors
// that follow see the positions a solver storing them in FP32 sees.
int oracle_set_f32_positions(void* h, int on) {
    ((Oracle*)h)->f32Positions = on != 0;
    return 0;
}

// The threshold of the DFSPH eps branches (|k_i + k_j| > eps and |k_i| >
// eps in the warm start and both solves, src/DFSPHSolver.cpp:26,364,368,
// 548,552,642,656,670; the alpha denominator test at :182).  Tests move it
// by a small relative amount to find the particles whose step depends on a
// branch decided within a solver's FP32 resolution of the threshold.
// computeDivergenceError (src/DFSPHSolver.cpp:408-492) of the current
// positions, neighbor lists and densities with the velocities `vel` (3n
// floats) in place of the oracle's own; the result replaces the divergence
// error field, the oracle's velocities are kept.  The divergence error is a
// cancelling sum of velocity differences, so a solver's value is checked
// both against the FP64 run's and against this, the same function of the
// velocities the solver holds.
int oracle_divergence_error(void* h, const float* vel) {
    Oracle* o = (Oracle*)h;
    const std::vector<V3> keep = o->v;
    for (size_t i = 0; i < o->N(); ++i) o->v[i] = V3(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
    o->computeDivergenceError();
    o->v = keep;
    return 0;
}

int oracle_set_eps(void* h, double eps) {
    ((Oracle*)h)->eps = eps;
    return 0;
}

int oracle_step(void* h, unsigned nsteps) {
    Oracle* o = (Oracle*)h;
    for (unsigned s = 0; s < nsteps; ++s) o->step();
    return 0;
}

unsigned oracle_num_fluid(void* h) { return (unsigned)((Oracle*)h)->N(); }
unsigned oracle_num_boundary(void* h) { return (unsigned)((Oracle*)h)->xb.size(); }

int oracle_get(void* h, int field, double* out) {
    Oracle* o = (Oracle*)h;
    switch (field) {
        case OF_X: put3(o->x, out); break;
        case OF_V: put3(o->v, out); break;
        case OF_A: put3(o->a, out); break;
        case OF_RHO: std::copy(o->rho.begin(), o->rho.end(), out); break;
        case OF_P: std::copy(o->p.begin(), o->p.end(), out); break;
        case OF_ALPHA: std::copy(o->alpha.begin(), o->alpha.end(), out); break;
        case OF_RHOSTAR: std::copy(o->rhoStar.begin(), o->rhoStar.end(), out); break;
        case OF_DIVERR: std::copy(o->divErr.begin(), o->divErr.end(), out); break;
        case OF_K: std::copy(o->k.begin(), o->k.end(), out); break;
        case OF_KV: std::copy(o->kv.begin(), o->kv.end(), out); break;
        case OF_BVOL: std::copy(o->vb.begin(), o->vb.end(), out); break;
        case OF_NORMAL: put3(o->normal, out); break;
        case OF_PACC: put3(o->pacc, out); break;
        case OF_ACC_STEP: put3(o->accStep, out); break;
        case OF_BX: put3(o->xb, out); break;
        case OF_PBN: put3(o->pbn, out); break;
        case OF_BDENS: std::copy(o->bdens.begin(), o->bdens.end(), out); break;
        case OF_BPRESS: std::copy(o->bpress.begin(), o->bpress.end(), out); break;
        case OF_PREDR: put3(o->predR, out); break;
        default: return -1;
    }
    return 0;
}

// Scalars: time, dt, iterations, iterationsV, avgDensError (last),
// avgDivError (last), maxDensityError, steps, sumIterations, sumIterationsV,
// maxVel, PCISPH delta, CFL time step of the last update, initial density
// error of the last step, initial divergence error of the last step.
int oracle_scalars(void* h, double* out) {
    Oracle* o = (Oracle*)h;
    out[0] = o->time; out[1] = o->ts; out[2] = o->iterations; out[3] = o->iterationsV;
    out[4] = o->avgDensError; out[5] = o->avgDivError; out[6] = o->maxDensityError;
    out[7] = o->steps; out[8] = o->sumIt; out[9] = o->sumItV; out[10] = o->maxVel;
    out[11] = o->delta; out[12] = o->lastCfl; out[13] = o->avgDensError0; out[14] = o->avgDivError0;
    return 0;
}

// Reference cell keys floor(x/h) of every fluid particle (n*3 ints).
int oracle_fluid_keys(void* h, int* out) {
    Oracle* o = (Oracle*)h;
    for (size_t i = 0; i < o->N(); ++i) {
        Key k = key_of(o->x[i], o->h);
        out[3 * i] = k.x; out[3 * i + 1] = k.y; out[3 * i + 2] = k.z;
    }
    return 0;
}

// Members of particle i's own cell in bucket order (ascending id, as
// inserted), and its candidate list (the cell's 27-cell concatenation).
// Returns the count written (at most cap), or -1.
int oracle_cell_points(void* h, unsigned i, unsigned* out, unsigned cap) {
    Oracle* o = (Oracle*)h;
    if (i >= o->N()) return -1;
    const Bucket& c = o->cellOf(o->x[i]);
    unsigned n = 0;
    for (const PointInfo& p : c.points) { if (n < cap) out[n] = p.id; ++n; }
    return (int)n;
}
int oracle_candidates(void* h, unsigned i, unsigned* out, unsigned cap, int boundary) {
    Oracle* o = (Oracle*)h;
    if (i >= o->N()) return -1;
    const Bucket& c = o->cellOf(o->x[i]);
    const std::vector<PointInfo>& src = boundary ? c.bneighbors : c.neighbors;
    unsigned n = 0;
    for (const PointInfo& p : src) { if (n < cap) out[n] = p.id; ++n; }
    return (int)n;
}

// ---- samplers (set-up path; the GPU library has its own implementation,
// compared bit-exactly in the tests)

// src/AkinciBoundayModel.cpp:100-134: hollow box shell, n = ceil(L/(h/2))
// intervals per axis.  Returns the count; writes up to cap points.
long oracle_sample_cube(const double* mn, const double* mx, double supportRadius, double* out, long cap) {
    double lx = std::fabs(mn[0] - mx[0]), ly = std::fabs(mn[1] - mx[1]), lz = std::fabs(mn[2] - mx[2]);
    unsigned nx = (unsigned)std::ceil(lx / (0.5 * supportRadius));
    unsigned ny = (unsigned)std::ceil(ly / (0.5 * supportRadius));
    unsigned nz = (unsigned)std::ceil(lz / (0.5 * supportRadius));
    double dx = lx / nx, dy = ly / ny, dz = lz / nz;
    long c = 0;
    for (unsigned i = 0; i <= nx; i++)
        for (unsigned j = 0; j <= ny; j++)
            for (unsigned k = 0; k <= nz; k++)
                if ((i == 0 || i == nx) || (j == 0 || j == ny) || (k == 0 || k == nz)) {
                    if (c < cap) {
                        out[3 * c] = i * dx + mn[0];
                        out[3 * c + 1] = j * dy + mn[1];
                        out[3 * c + 2] = k * dz + mn[2];
                    }
                    ++c;
                }
    return c;
}

// src/PCISPHBoundaryModel.cpp:60-189: the box surface at ceil(l / 0.45h)
// intervals per axis, each point with the normal of its corner, edge or face
// (faces: +x at i = 0, -x at i = nx, +y at j = 0, -y at j = ny, +z at k = 0,
// -z at k = nz; edges and corners glm::normalize of the mean), tested in the
// reference's order; times -1 unless inverted (:7-27).
static V3 pcisph_box_normal(unsigned i, unsigned j, unsigned k, unsigned nx, unsigned ny, unsigned nz) {
    const V3 L(1, 0, 0), R(-1, 0, 0), T(0, -1, 0), Bo(0, 1, 0), F(0, 0, -1), Be(0, 0, 1);
    auto nrm = [](V3 v) { return v * (1.0 / std::sqrt(dot(v, v))); };
    auto e2 = [&](V3 a, V3 b) { return nrm((a + b) * 0.5); };
    auto e3 = [&](V3 a, V3 b, V3 c) { return nrm(((a + b) + c) / 3.0); };
    const bool i0 = i == 0, iN = i == nx, im = i > 0 && i < nx;
    const bool j0 = j == 0, jN = j == ny, jm = j > 0 && j < ny;
    const bool k0 = k == 0, kN = k == nz, km = k > 0 && k < nz;
    if (i0 && j0 && k0) return e3(L, Be, Bo);
    if (iN && j0 && k0) return e3(R, Be, Bo);
    if (i0 && j0 && kN) return e3(L, F, Bo);
    if (iN && j0 && kN) return e3(R, F, Bo);
    if (i0 && jN && k0) return e3(L, Be, T);
    if (iN && jN && k0) return e3(R, Be, T);
    if (i0 && jN && kN) return e3(L, F, T);
    if (iN && jN && kN) return e3(R, F, T);
    if (i0 && j0 && km) return e2(L, Bo);
    if (im && j0 && k0) return e2(Be, Bo);
    if (im && j0 && kN) return e2(F, Bo);
    if (iN && j0 && km) return e2(R, Bo);
    if (i0 && jN && km) return e2(L, T);
    if (im && jN && k0) return e2(Be, T);
    if (im && jN && kN) return e2(F, T);
    if (iN && jN && km) return e2(R, T);
    if (i0 && jm && k0) return e2(L, Be);
    if (iN && jm && k0) return e2(R, Be);
    if (iN && jm && kN) return e2(R, F);
    if (i0 && jm && kN) return e2(L, F);
    if (i0 && jm && km) return L;
    if (iN && jm && km) return R;
    if (im && jN && km) return T;
    if (im && j0 && km) return Bo;
    if (im && jm && kN) return F;
    return Be;   // im && jm && k0
}
long oracle_sample_box_normals(const double* mn, const double* mx, double supportRadius, int inverted, double* out,
                               double* nout, long cap) {
    double lx = std::fabs(mn[0] - mx[0]), ly = std::fabs(mn[1] - mx[1]), lz = std::fabs(mn[2] - mx[2]);
    unsigned nx = (unsigned)std::ceil(lx / (0.45 * supportRadius));
    unsigned ny = (unsigned)std::ceil(ly / (0.45 * supportRadius));
    unsigned nz = (unsigned)std::ceil(lz / (0.45 * supportRadius));
    double dx = lx / nx, dy = ly / ny, dz = lz / nz;
    const double inv = inverted ? 1.0 : -1.0;
    long c = 0;
    for (unsigned i = 0; i <= nx; i++)
        for (unsigned j = 0; j <= ny; j++)
            for (unsigned k = 0; k <= nz; k++)
                if ((i == 0 || i == nx) || (j == 0 || j == ny) || (k == 0 || k == nz)) {
                    if (c < cap) {
                        out[3 * c] = i * dx + mn[0];
                        out[3 * c + 1] = j * dy + mn[1];
                        out[3 * c + 2] = k * dz + mn[2];
                        const V3 nn = inv * pcisph_box_normal(i, j, k, nx, ny, nz);
                        nout[3 * c] = nn.x;
                        nout[3 * c + 1] = nn.y;
                        nout[3 * c + 2] = nn.z;
                    }
                    ++c;
                }
    return c;
}

// src/AkinciBoundayModel.cpp:136-164: equal-area rings.
long oracle_sample_sphere(const double* origin, double radius, double particleRadius, double* out, long cap) {
    double separacion = 1.5;
    unsigned num_parts = (unsigned)std::floor(4.0 * M_PI * radius * radius / (M_PI * particleRadius * particleRadius) * separacion);
    double a = 4.0 * M_PI / num_parts;
    double dd = std::sqrt(a);
    int m_theta = int(std::round(M_PI / dd));
    double d_theta = M_PI / m_theta;
    double d_phi = a / d_theta;
    long c = 0;
    for (int mm = 0; mm < m_theta; ++mm) {
        double theta = M_PI * (mm + 0.5) / m_theta;
        int m_phi = int(std::round(2.0 * M_PI * std::sin(theta) / d_phi));
        for (int nn = 0; nn < m_phi; ++nn) {
            double phi = 2.0 * M_PI * nn / m_phi;
            double px = std::sin(theta) * std::cos(phi);
            double py = std::sin(theta) * std::sin(phi);
            double pz = std::cos(theta);
            if (c < cap) {
                out[3 * c] = radius * px + origin[0];
                out[3 * c + 1] = radius * py + origin[1];
                out[3 * c + 2] = radius * pz + origin[2];
            }
            ++c;
        }
    }
    return c;
}

// src/Simulation.cpp:512-539: lattice min + r + (i,j,k)*2r, per-edge count
// floor(L/2r + 1e-5), i outermost.
long oracle_fluid_block(const double* mn, const double* mx, double particleRadius, double* out, long cap) {
    double dist = 2.0 * particleRadius;
    int px = (int)std::floor((mx[0] - mn[0]) / dist + 1e-5);
    int py = (int)std::floor((mx[1] - mn[1]) / dist + 1e-5);
    int pz = (int)std::floor((mx[2] - mn[2]) / dist + 1e-5);
    long c = 0;
    for (int i = 0; i < px; ++i)
        for (int j = 0; j < py; ++j)
            for (int k = 0; k < pz; ++k) {
                if (c < cap) {
                    out[3 * c] = (double)i * dist + (mn[0] + particleRadius);
                    out[3 * c + 1] = (double)j * dist + (mn[1] + particleRadius);
                    out[3 * c + 2] = (double)k * dist + (mn[2] + particleRadius);
                }
                ++c;
            }
    return c;
}

// Single kernel evaluations for closed-form checks.
double oracle_kernel_W(double h, double r) { Kernels k; k.set(h); return k.W(r); }
void oracle_kernel_gradW(double h, const double* r, double* out) {
    Kernels k; k.set(h); V3 g = k.gradW(V3(r[0], r[1], r[2])); out[0] = g.x; out[1] = g.y; out[2] = g.z;
}
double oracle_kernel_cohesion(double h, double r) { Kernels k; k.set(h); return k.cohesion(V3(r, 0, 0)); }
double oracle_kernel_adhesion(double h, double r) { Kernels k; k.set(h); return k.adhesion(V3(r, 0, 0)); }
void oracle_kernel_spiky_grad(double h, const double* r, double* out) {
    Kernels k; k.set(h); V3 g = k.spikyGrad(V3(r[0], r[1], r[2])); out[0] = g.x; out[1] = g.y; out[2] = g.z;
}
void oracle_kernel_poly6_grad(double h, const double* r, double* out) {
    Kernels k; k.set(h); V3 g = k.poly6Grad(V3(r[0], r[1], r[2])); out[0] = g.x; out[1] = g.y; out[2] = g.z;
}

}  // extern "C"
