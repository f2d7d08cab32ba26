// sphg_internal.h -- device-side data layout shared by the HIP kernels and
// the C-ABI layer of libsphg.
//
// One scene = one fluid model + one Akinci boundary model (the reference's
// own constraint, src/Simulation.h:24-29 and src/AkinciBoundayModel.cpp:32).
// Every kernel walks a chunk table: each 256-thread block owns one chunk of
// one scene, so a launch covers any number of scenes of uneven size
// (batching) and a single scene is a batch of one.
//
// HBM layout per scene (FP32, struct-of-arrays, particles sorted by
// (cell, id) every step):
//   x, v          float4 (x.w: DFSPH v_x copy; v.w unused) -- double-buffered for the sort
//   vyz           float2 (v_y, v_z) copy (DFSPH), beside x.w for the error sweeps
//   xk            float4 (x, y, z, s)             -- s = the scalar the next
//                 neighbor sweep gathers (warm-start k, kappa, kappa_v, p/rho^2)
//   gb            float4 (sum_b rho0 V_b gradW_ib, sum_b rho0 V_b W_ib)
//   drho = rho - rho0, alpha, drhoS = rho* - rho0, divErr, p  float
//   k, kv         float (DFSPH warm-start accumulators, double-buffered)
//   id            uint32 original particle id (double-buffered)
//   cell          int    linear cell of each sorted particle
//   nbr           uint32 neighbor list inside 64-particle groups, in blocks of
//                 4 slots, lane-major within a block (a lane's 4 next slots
//                 are one 16-byte load)
//   nnb, ngrad    uint16/uint8 neighbor count and count of nonzero-gradient
//                 neighbors (the DFSPH 20-neighbor cutoff)
//   gcnt          uint16 per 64-particle group: the group's longest list,
//                 rounded up to 4; shorter lists are padded with the
//                 particle's own index, whose every term is exactly zero
// Boundary: xb float4 (x, y, z, V_b) sorted by cell, bCellStart table.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sphg {

// The block size and the scan chunk are compile-time choices of the layout
// (the chunk tables, the lists' groups and the reduction partials are built
// on them); a build may pick others with -DSPHG_BLOCK=128|256|512 and
// -DSPHG_CELL_CHUNK=<multiple of the block, up to 16 blocks>.
#ifndef SPHG_BLOCK
#define SPHG_BLOCK 256
#endif
#ifndef SPHG_CELL_CHUNK
#define SPHG_CELL_CHUNK 1024
#endif
constexpr int kBlock = SPHG_BLOCK;            // threads per block = particles per chunk
constexpr int kWaves = kBlock / 64;           // wavefronts per block
constexpr int kGroup = 64;                    // particles per neighbor-list group (one wavefront)
constexpr int kCellChunk = SPHG_CELL_CHUNK;   // cells per scan chunk (1024: ~300 blocks for a 1M-particle grid)
constexpr int kCellPerThread = kCellChunk / kBlock;
static_assert(kBlock == 128 || kBlock == 256 || kBlock == 512, "block of 2, 4 or 8 wavefronts");
static_assert(kCellChunk % kBlock == 0 && kCellPerThread >= 1 && kCellPerThread <= 16,
              "a scan chunk is 1 to 16 cells per thread");

enum Method { WCSPH = 0, PCISPH = 1, DFSPH = 2 };
enum Flags { F_XSPH = 1, F_ST = 2, F_ADH = 4, F_VSTD = 8, F_VART = 16 };
// any non-pressure force: the step's acceleration a is nonzero
constexpr int F_FORCES = F_XSPH | F_ST | F_ADH | F_VSTD | F_VART;

// FP32 kernel coefficients (computed in FP64 on the host, then rounded).
struct KCoef {
    float h, invh, h2;
    float ck;      // cubic spline W:    8 / (pi h^3)        src/CubicSpline.h:25
    float clh2;    // cubic grad, q<=1/2: 48/(pi h^3) / h^2  src/CubicSpline.h:26,64
    float clh;     // cubic grad, q>1/2:  48/(pi h^3) / h
    float spg;     // spiky grad:        45 / (pi h^6)      src/Spiky.h:26
    float co1;     // cohesion:          32 / (pi h^9)      src/Cohesion.h:26
    float co2;     // cohesion offset:   h^6 / 64           src/Cohesion.h:27
    float hlo2, hhi2;   // h^2 (1 -/+ 1e-4): FP32 pre-filter band of the exact test
};

// FP64 copies used by the neighbor sums (FP32 state, FP64 arithmetic).
struct KCoefD {
    double h, invh, h2;
    double ck, clh2, clh, spg, co1, co2;
    double hlo2, hhi2;   // h^2 (1 -/+ 1e-9): band of the operation-exact neighbor test
    double gA, gB, hHalf;  // cubic gradient for q <= 1/2 as gA r + gB; h/2
    double vl;             // viscosity Laplacian 45/(pi h^6) (src/ViscoK.h:23)
    double p6v;            // Poly6 value coefficient 315/(64 pi h^9) (src/Poly6.h:27)
};

// Device-resident control state of one scene (one writer at a time: the
// last block of a reduction, or the first block of a scene).
struct Scalars {
    double dt, time, maxVel, cflTs;
    double avgErr, avgErr0, avgErrV, avgErrV0, maxDensErr;
    int it, itV, done, doneV;
    unsigned steps, sumIt, sumItV, ctr;
    unsigned overflow;          // largest neighbor count above max_neighbors in the lists the current
                                // step call has used (0: none)
    unsigned listOver;          // the same for the lists built last (reset by the key pass before each build)
    unsigned nActive;           // active particles (the emitters raise it)
    unsigned shellBad;          // PCISPH: this step's shell lists cannot stand in for the 27-cell walk
    unsigned maxD2, maxD2Prev;  // PCISPH: largest squared predicted displacement (FP32 bits) of this
                                // step so far and of the previous step, which sizes this step's shell
    unsigned shellSteps, walkSteps;  // PCISPH: steps whose predicted densities used the shell lists /
                                     // took the 27-cell walk for every particle
    unsigned d0Valid;           // DFSPH: d0 holds the divergence of the current velocities (written by
                                // the list build and the divergence error sweeps; cleared by the cube
                                // clamp, which changes the velocities after them)
};

// One emitter's constants and state (src/Emitter.h:9-55).  Its group's
// local positions are rows [group, group + groupSize) of the scene's group
// table (3 doubles each).
struct EmitterDev {
    double pos[3];              // emitter origin
    double q[4];                // rotation quaternion (w, x, y, z)
    double vel;                 // speed along local +z
    double interval;            // spacing 1.105 2r / vel between groups
    double next;                // next emission time
    unsigned remaining;         // budget left
    int group, groupSize;
    int pad;
};

struct DevScene {
    KCoef kc;
    KCoefD kd;
    double hD;                  // support radius in FP64 (cell keys, exact neighbor test)
    double r, cfl, dtMin, dtMax, eta, etaV;
    float m, rho0, W0, pciDelta;
    double mD, rho0D, pciDeltaD;
    float gx, gy, gz, pad0;
    float nu, nub, stCoef, adhBeta;
    float stiffness, gammaEos, pad1, pad2;
    int method, flags;
    int cubeOn;                 // CubeBoundaryModel clamp after the step's integration
    // box faces as the nearest FP32 values inside the box (min rounded up,
    // max rounded down): for an FP32 coordinate p, p < cubeLo is the same
    // test as p < min in FP64, and a clamped particle stays inside
    float cubeLo[3], cubeHi[3];
    double cubeNf, cubeTf;
    int minIt, maxIt, minItV, maxItV;
    int kx0, ky0, kz0;          // key of the first unpadded cell
    int nx, ny, nz, ncells;     // padded grid (one empty layer on every side)
    int n, nb, maxNbr, nChunks;
    int nCellChunks, cap;       // cap: particle slots (active + every emitter's remaining budget)
    int nEmit, pad3;
    EmitterDev *emit;
    const double *emitGroup;
    // PCISPH boundary method (src/PCISPHBoundaryModel.h:10-52): boundary
    // particles carry normals, densities and pressures (xb.w = (2r)^3, so
    // rho0 xb.w is the boundary mass); the fluid carries the normal the
    // collision projection last wrote (NaN where nothing collided)
    int pbOn, pad4;
    double pbNf, pbTf;
    float2 *vyz;                // DFSPH: (v_y, v_z), written by the velocity sweeps with v_x in x.w;
                                // the error sweeps gather x and vyz, 24 B per pair.  In the density
                                // solve the two hold u, the step's velocity change so far, instead
    double *d0;                 // DFSPH: sum_j m (v_i - v_j).gradW_ij + v_i.gb_i of the current
                                // velocities: the density solve's constant part of the predicted
                                // density sum, from the velocity at the step's start
    int pciShell, pad5;         // PCISPH: predicted densities from the list and a shell list
    uint16_t *nsh;              // PCISPH: shell entries per particle (candidates within h + skin
                                // that are not in the list, stored from the list's last slot down)
    float4 *bn;                 // boundary normals (sorted like xb)
    double *bself;              // per boundary particle: boundary sums of m_b W (cubic) and m_b Poly6
    float *bdens, *bpres, *bpr; // boundary density, pressure, pressure / density^2
    float4 *pbn, *pbn2;         // fluid normals; DFSPH carries them through the sort
    float4 *pv;                 // PCISPH predicted velocity of the next density sweep
    float4 *prS, *pvS, *x0s;    // PCISPH: last iteration's corrected displacement and predicted
                                // velocity; positions before the integration
    uint8_t *everOcc;           // per cell: a particle was ever inserted there
    // particle arrays; the *2 members are the other half of each double buffer
    float4 *x, *x2, *v, *v2;
    float *k, *k2, *kv, *kv2;
    uint32_t *id, *id2;
    float4 *a, *xk, *gb, *aux, *xs, *nrm, *pacc, *gbs, *adh;
    float *drho;                // rho - rho0 (keeps the deviation's precision in FP32)
    float *alpha, *drhoS, *divErr, *p;   // drhoS = rho* - rho0
    int *cellU;                 // cell of each particle before the sort
    int *cell;                  // cell of each particle after the sort
    uint32_t *perm;
    int *cellCount, *cellStart, *cellChunkSum;
    uint32_t *nbr;
    uint16_t *nnb;
    uint8_t *ngrad;
    uint16_t *gcnt;             // per 64-particle group: list length padded to a multiple of 4
    float4 *xb;
    int *bCellStart;
    double *partial;            // one per particle chunk
    Scalars *ss;
};

struct Chunk {
    int scene;   // index into the launch's scene table
    int begin;   // first particle (or first cell) of the chunk inside the scene
};

}  // namespace sphg
