/*
 * sphg.h -- C ABI of libsphg, the gfx950 (MI355X) SPH solver hot path.
 *
 * The path behind this ABI is the reference's SPHSolver::step for WCSPH,
 * PCISPH and DFSPH with Akinci boundaries, XSPH viscosity, Akinci surface
 * tension and Akinci adhesion (josemanuel28v/SPH src/SPHSolver.h:6-50,
 * src/WCSPHSolver.cpp, src/PCISPHSolver.cpp, src/DFSPHSolver.cpp).  The C++
 * host layer (sph_amd/host: Simulation, GpuSolver, FluidModel, ...) calls
 * these functions where the reference calls its OpenMP solver.
 *
 * Conventions: every function returns 0 (SPHG_OK) or a negative SPHG_E*
 * status and never throws; sphg_last_error() gives the message of the
 * calling thread's last failure.  Arrays are plain host pointers, borrowed
 * for the duration of the call and copied before it returns; xyz arrays hold
 * n*3 floats in particle-id order.  A context is used by one host thread at
 * a time, with one exception: sphg_scene_create, the set_fluid /
 * set_boundary / set_boundary_normals / add_emitter calls and
 * sphg_scene_plan_bytes on a scene that has not been initialised, and
 * sphg_scene_destroy of such a scene, touch host memory only and may run on
 * another thread while the context runs.  All device memory is owned by the
 * library.
 */
#ifndef SPHG_H
#define SPHG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    SPHG_OK = 0,
    SPHG_EINVAL = -1,    /* bad argument or scene description */
    SPHG_EHIP = -2,      /* HIP runtime error */
    SPHG_ENOMEM = -3,    /* device allocation failed */
    SPHG_ESTATE = -4,    /* call out of order (e.g. step before init) */
    SPHG_EOVERFLOW = -5  /* a particle had more neighbors than max_neighbors */
};

/* Solver and boundary method numbers of the reference scene format
 * (src/Simulation.cpp:21-33). */
enum { SPHG_WCSPH = 0, SPHG_PCISPH = 1, SPHG_DFSPH = 2 };
/* boundaryMethod (src/Simulation.cpp:25-27): cube (a clamp into a box, or
 * nothing when the scene has no CubeBoundaryModel), Akinci particles.  The
 * PCISPH boundary method (1) is not available. */
enum { SPHG_BOUNDARY_CUBE = 0, SPHG_BOUNDARY_PCISPH = 1, SPHG_BOUNDARY_AKINCI = 2 };
/* viscosityMethod (src/Simulation.h enums): standard SPH (src/StandardViscosityForce.h),
 * artificial (src/ArtificialViscosity.h), XSPH (src/XSPHViscosity.h) */
enum { SPHG_VISCOSITY_STANDARD = 0, SPHG_VISCOSITY_ARTIFICIAL = 1, SPHG_VISCOSITY_XSPH = 2 };
enum { SPHG_SURFACE_TENSION_AKINCI = 3, SPHG_ADHESION_AKINCI = 4 };

typedef struct sphg_ctx sphg_ctx;
typedef struct sphg_scene sphg_scene;
typedef struct sphg_batch sphg_batch;

/* One scene: the reference's SceneInfo/FluidInfo fields
 * (src/Simulation.h:47-68, read at src/SceneLoader.cpp:38-141) plus the
 * device-side sizing.  Replaces the solver set-up done by
 * Simulation::importScene (src/Simulation.cpp:342-424). */
typedef struct sphg_scene_desc {
    int32_t method;                 /* SPHG_WCSPH / SPHG_PCISPH / SPHG_DFSPH */
    int32_t boundary_method;        /* SPHG_BOUNDARY_AKINCI, or SPHG_BOUNDARY_CUBE (no boundary particles) */
    double particle_radius;         /* r; support radius h = 4r (src/Simulation.cpp:197) */
    double density0;                /* rest density rho0 */
    double gravity[3];
    double time_step;               /* initial (DFSPH) or fixed (WCSPH, PCISPH) dt */
    double min_time_step, max_time_step, fps;   /* dt clamps; min = min(1/fps, min) */
    double cfl_factor;              /* DFSPH */
    double stiffness, gamma;        /* WCSPH Tait EOS */
    double eta, eta_v;              /* PCISPH/DFSPH density, DFSPH divergence tolerance */
    uint32_t min_iterations, max_iterations, min_iterations_v, max_iterations_v;
    int32_t viscosity_method;       /* -1 or SPHG_VISCOSITY_STANDARD / _ARTIFICIAL / _XSPH */
    double viscosity, boundary_viscosity;
    int32_t surface_tension_method; /* -1 or SPHG_SURFACE_TENSION_AKINCI */
    double surface_tension;
    int32_t adhesion_method;        /* -1 or SPHG_ADHESION_AKINCI */
    double adhesion;
    uint32_t max_neighbors;         /* neighbor-list capacity; 0 = 96 */
    double domain_min[3], domain_max[3];  /* grid bounds; all zero = fit the particles */
    /* CubeBoundaryModel (boundary_method SPHG_BOUNDARY_CUBE,
     * src/CubeBoundaryModel.cpp:5-94): when cube_enabled, every fluid
     * position is clamped into [cube_min, cube_max] after the step's
     * integration, the velocity component leaving the box is multiplied by
     * -cube_normal_fct and the others by cube_tangential_fct (defaults 0, 1). */
    int32_t cube_enabled;
    double cube_min[3], cube_max[3];
    double cube_normal_fct, cube_tangential_fct;
    /* PCISPHBoundaryModel (boundary_method SPHG_BOUNDARY_PCISPH,
     * src/PCISPHBoundaryModel.h:21-22; the scene file's Boundary[].normalFct
     * and tangFct): the velocity of a fluid particle projected out of the
     * boundary becomes tangential_fct * tangential - normal_fct * normal. */
    double pb_normal_fct, pb_tangential_fct;
} sphg_scene_desc;

/* Emitter types (src/Emitter.cpp:7-8). */
enum { SPHG_EMITTER_SQUARE = 0, SPHG_EMITTER_CIRCLE = 1 };

/* One particle emitter (src/Emitter.h:9-55), with the scene file's emitter
 * keys (src/SceneLoader.cpp:123-139).  Every time the simulation time
 * reaches the next emission time and a whole group is left in the budget,
 * one group (a square lattice of width x height, or a disc of radius width,
 * at spacing 2r) enters with velocity `velocity` along the emitter's local
 * +z, rotated by the quaternion and offset to `position`; the next emission
 * time advances by spacing * 1.105 * 2r / velocity. */
typedef struct sphg_emitter_desc {
    uint32_t type;                  /* SPHG_EMITTER_SQUARE or SPHG_EMITTER_CIRCLE */
    uint32_t num_particles;         /* budget: particles emitted in total */
    double position[3];
    double velocity;
    double rotation[4];             /* quaternion (w, x, y, z), the scene file's order */
    double start_time;              /* first emission time */
    double width, height;           /* square sides; a circle's radius is width */
    double spacing;
} sphg_emitter_desc;

/* Device-side per-scene state, as the reference's solver members
 * (src/DFSPHSolver.h:26-31, src/PCISPHSolver.h:17, src/TimeManager.h:30-35). */
typedef struct sphg_scene_scalars {
    double time, time_step, max_vel, cfl_time_step;
    double avg_density_error;       /* last density-solver iteration */
    double avg_density_error0;      /* before the density loop */
    double avg_divergence_error, avg_divergence_error0;
    double max_density_error;       /* PCISPH */
    int32_t iterations, iterations_v;
    uint32_t steps, sum_iterations, sum_iterations_v;
    uint32_t neighbor_overflow;     /* largest neighbor count above max_neighbors during the last
                                       step / batch call that stepped this scene, or 0 */
    /* PCISPH predicted densities: steps that read the neighbor and shell
     * lists (particles whose shell did not fit walk their 27 cells; see
     * SPHG_FIELD_NSH), and steps whose displacement guard sent every particle
     * to the 27-cell walk (always, with the PCISPH boundary method) */
    uint32_t pcisph_shell_steps, pcisph_walk_steps;
} sphg_scene_scalars;

typedef struct sphg_step_stats {
    double gpu_ms;                  /* device time of the stepped region (HIP events) */
    uint32_t steps;
    uint32_t iterations;            /* density/pressure iterations summed over steps */
    uint32_t iterations_v;          /* DFSPH divergence iterations summed over steps */
    uint32_t host_syncs;
} sphg_step_stats;

/* Fields for sphg_scene_get, returned in particle-id order. */
enum {
    SPHG_FIELD_X = 0,        /* float[n*3] positions */
    SPHG_FIELD_V = 1,        /* float[n*3] velocities */
    SPHG_FIELD_A = 2,        /* float[n*3] acceleration of the last step before gravity and pressure
                                correction (WCSPH: pressure + non-pressure forces) */
    SPHG_FIELD_RHO = 3,      /* float[n] density (PCISPH: last predicted density) */
    SPHG_FIELD_P = 4,        /* float[n] pressure (WCSPH, PCISPH) */
    SPHG_FIELD_ALPHA = 5,    /* float[n] DFSPH alpha */
    SPHG_FIELD_RHOSTAR = 6,  /* float[n] DFSPH predicted density of the last density iteration */
    SPHG_FIELD_DIVERR = 7,   /* float[n] DFSPH divergence error of the last divergence iteration */
    SPHG_FIELD_K = 8,        /* float[n] DFSPH density warm-start accumulator */
    SPHG_FIELD_KV = 9,       /* float[n] DFSPH divergence warm-start accumulator */
    SPHG_FIELD_BVOL = 10,    /* float[nb] Akinci boundary volumes, boundary input order */
    SPHG_FIELD_NORMAL = 11,  /* float[n*3] surface-tension normals */
    SPHG_FIELD_PACC = 12,    /* float[n*3] PCISPH pressure acceleration */
    SPHG_FIELD_KEYS = 13,    /* int32[n*3] cell keys floor(x/h) */
    SPHG_FIELD_NNB = 14,     /* int32[n] fluid neighbors with |r| <= h, self excluded */
    SPHG_FIELD_NGRAD = 15,   /* int32[n] of which with nonzero kernel gradient (saturates at 255; the
                                divergence error's cutoff is at 20) */
    SPHG_FIELD_PBNORMAL = 16,  /* float[n*3] PCISPH boundary method: the fluid normal of the last
                                  collision projection (FluidModel::n; NaN where nothing collided) */
    SPHG_FIELD_BDENSITY = 17,  /* float[nb] PCISPH boundary method: boundary densities, input order */
    SPHG_FIELD_BPRESSURE = 18, /* float[nb] PCISPH boundary method: boundary pressures, input order */
    SPHG_FIELD_DRHO = 19,      /* float[n] rho - rho0 as stored (the density's deviation at full FP32
                                  precision; FIELD_RHO adds rho0 and rounds again) */
    SPHG_FIELD_DRHOSTAR = 20,  /* float[n] DFSPH rho* - rho0 as stored, the error the density solve runs on */
    SPHG_FIELD_NSH = 21        /* int32[n] PCISPH: shell-list entries of each particle at the last list
                                  build, -1 where the shell did not fit (that particle walks its cells) */
};

const char* sphg_version(void);
const char* sphg_last_error(void);

/* Contexts: one device, one stream. */
int sphg_ctx_create(int32_t device, sphg_ctx** out);
int sphg_ctx_destroy(sphg_ctx* ctx);
int sphg_ctx_synchronize(sphg_ctx* ctx);
/* Replay of captured step segments (default off).  On, the fixed launch
 * sequences of a step (everything but the host-decided iterations of the
 * pressure loops) are captured once into HIP graphs per scene / per
 * persistent batch and replayed; off, every kernel is launched one by one.
 * Results are identical either way; kernel timing launches one by one.  No
 * counterpart in the reference (its substep is a sequence of OpenMP loops,
 * src/DFSPHSolver.cpp:33-87). */
int sphg_ctx_graphs(sphg_ctx* ctx, int32_t enable);

/* Scenes.  create -> set_boundary / set_fluid -> init -> step ...
 * sphg_scene_create replaces Simulation::setSimulationMethod and the solver
 * setters (src/Simulation.cpp:207-231, 342-380). */
int sphg_scene_create(sphg_ctx* ctx, const sphg_scene_desc* desc, sphg_scene** out);
int sphg_scene_destroy(sphg_scene* scene);
/* FluidModel::init(points, velocities) (src/FluidModel.cpp:37-49); vel may be NULL.
 * A scene holds at most 2^28 particle slots (fluid plus emitter budgets),
 * and slots x max_neighbors must stay below 2^32 (44.7M slots at the
 * default 96): sphg_scene_init / plan_bytes refuse larger ones. */
int sphg_scene_set_fluid(sphg_scene* scene, const float* xyz, const float* vel, uint32_t n);
/* BoundaryModel::init / AkinciBoundaryModel::addCube... (src/BoundaryModel.cpp:3-9). */
int sphg_scene_set_boundary(sphg_scene* scene, const float* xyz, uint32_t nb);
/* PCISPHBoundaryModel normals (src/PCISPHBoundaryModel.cpp:7-52): one per
 * boundary particle set by set_boundary, required for boundary_method
 * SPHG_BOUNDARY_PCISPH; see sphg_sample_box_normals / _sphere_normals. */
int sphg_scene_set_boundary_normals(sphg_scene* scene, const float* nrm, uint32_t nb);
/* Replace the fluid normal field (id order, n*3) of the PCISPH boundary
 * method, e.g. with an injected state: DFSPH's velocity correction at the
 * start of a step reads the normals of the last step's projection. */
int sphg_scene_set_pb_normals(sphg_scene* scene, const float* nrm);
/* FluidModel::addEmitter (src/FluidModel.cpp:124-143): after set_fluid (or
 * with no fluid set: an inflow-only scene), before init.  The capacity
 * grows by the emitter's budget; emitted particles are activated on the
 * device, in emitter order, with ids continuing from the active count
 * (Emitter::emitParticles, src/Emitter.cpp:85-162). */
int sphg_scene_add_emitter(sphg_scene* scene, const sphg_emitter_desc* emitter);
/* FluidModel::getNumParticles: fluid particles plus every emitter's budget.
 * sphg_scene_info's n is the active count (getNumActiveParticles), which the
 * field read-backs cover. */
int sphg_scene_capacity(sphg_scene* scene, uint32_t* capacity);
/* Simulation::init -> solver->init(): boundary volumes, and for DFSPH the
 * first grid, densities and alpha (src/Simulation.cpp:102-110,
 * src/SPHSolver.cpp:22-39, src/DFSPHSolver.cpp:24-36). */
int sphg_scene_init(sphg_scene* scene);
/* Replace the particle state (id order) and rebuild what derives from the
 * positions, as after a DFSPH step's grid rebuild (src/DFSPHSolver.cpp:
 * 87-97).  k, kv may be NULL (zero).  last_iterations feeds the DFSPH dt
 * rule (src/DFSPHSolver.cpp:857-860). */
int sphg_scene_set_state(sphg_scene* scene, const float* xyz, const float* vel, const float* k, const float* kv,
                         double time_step, uint32_t last_iterations);
/* nsteps x SPHSolver::step() (src/SPHSolver.h:45; WCSPH src/WCSPHSolver.cpp:
 * 14-78, PCISPH src/PCISPHSolver.cpp:21-81, DFSPH src/DFSPHSolver.cpp:38-112).
 * stats may be NULL. */
int sphg_scene_step(sphg_scene* scene, uint32_t nsteps, sphg_step_stats* stats);
/* Field read-back in id order (FluidModel getters, src/FluidModel.h:56-66;
 * DFSPHSolver getters src/DFSPHSolver.h:60-62). */
int sphg_scene_get(sphg_scene* scene, int32_t field, void* out);
int sphg_scene_get_scalars(sphg_scene* scene, sphg_scene_scalars* out);
/* Grid parity read-back: ids in sorted (cell, id) order and their linear
 * cells; both arrays of n.  Buckets of the reference's HashTable
 * (src/HashTable.h:45-53) are the runs of equal cell. */
int sphg_scene_get_sorted(sphg_scene* scene, uint32_t* ids, int32_t* cells);
/* Neighbor lists in id order: counts[n], ids[n*cap] (row i lists the ids of
 * i's neighbors, unordered).  cap >= max_neighbors. */
int sphg_scene_get_neighbors(sphg_scene* scene, uint32_t* counts, uint32_t* ids, uint32_t cap);
/* Asynchronous frames (SceneLoader::writeFluid, src/SceneLoader.cpp:204-268:
 * positions and velocities of the active particles in id order, with the
 * time and time step).  A snapshot object owns a device staging buffer and
 * pinned host buffers sized for the scene's capacity.  _begin enqueues the
 * read-back on the scene's stream and returns at once: the copy to host
 * memory runs on a second stream while later steps compute.  _wait blocks
 * until that copy is complete and returns pointers into the snapshot's host
 * buffers (valid until its next _begin or _destroy); it may be called from
 * another thread than the one driving the context.  A snapshot's next
 * _begin waits on the device for its previous copy.  No reference
 * counterpart for the split: the reference writes frames synchronously. */
typedef struct sphg_snapshot sphg_snapshot;
int sphg_snapshot_create(sphg_scene* scene, sphg_snapshot** out);
int sphg_snapshot_begin(sphg_snapshot* snap);
int sphg_snapshot_wait(sphg_snapshot* snap, const float** xyz, const float** vel, uint32_t* n, double* time,
                       double* time_step);
int sphg_snapshot_destroy(sphg_snapshot* snap);

/* Sizes: fluid, boundary, grid cells, device bytes held. */
int sphg_scene_info(sphg_scene* scene, uint32_t* n, uint32_t* nb, uint32_t* ncells, uint64_t* device_bytes);

/* Step several scenes of one method together: one launch per solver phase
 * covers every scene (no reference counterpart: the reference runs one
 * scene per process through its Simulation singleton,
 * src/Simulation.cpp:84-90). */
int sphg_batch_step(sphg_ctx* ctx, sphg_scene* const* scenes, uint32_t nscenes, uint32_t nsteps,
                    sphg_step_stats* stats);
/* sphg_batch_step without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, serving loops that forbid
 * hipMalloc).  No counterpart in the reference.  The batch's tables -- two
 * scene tables of 928 bytes per scene, 8 bytes per 256 fluid slots and 8
 * bytes per 1024 grid cells, each rounded up to 256 bytes -- are written to
 * a static device array of SPHG_NOALLOC_SCRATCH_BYTES that each device holds
 * until the process ends; a batch that needs more is refused with
 * SPHG_EINVAL.  Every call on a device shares that array: calls that
 * overlap on one device (from two contexts or threads) race.  The scenes
 * themselves were allocated by sphg_scene_init; gpu_ms is the host clock
 * around the steps (no HIP events are created), and kernel timing must be
 * off.  Results are those of sphg_batch_step, bit for bit. */
#define SPHG_NOALLOC_SCRATCH_BYTES (8u << 20)
int sphg_batch_step_noalloc(sphg_ctx* ctx, sphg_scene* const* scenes, uint32_t nscenes, uint32_t nsteps,
                            sphg_step_stats* stats);
/* The same with the batch kept between calls (its chunk tables and its
 * iteration-count prediction), for a serving loop that steps a changing set
 * of resident scenes in quanta: create when the set changes, run per
 * quantum, destroy before any of its scenes.  No reference counterpart. */
int sphg_batch_create(sphg_ctx* ctx, sphg_scene* const* scenes, uint32_t nscenes, sphg_batch** out);
int sphg_batch_run(sphg_batch* batch, uint32_t nsteps, sphg_step_stats* stats);
int sphg_batch_destroy(sphg_batch* batch);
/* Launch-schedule autotuner of a scene's own stepping or of a persistent
 * batch.  The schedule is how many pressure-loop iterations are launched
 * past the previous step's count before the first convergence read-back
 * (density loop 0-3, DFSPH divergence loop 0-1) and whether the step's
 * fixed launch segments are replayed as HIP graphs.  Every schedule gives
 * the same results, so the trials are ordinary steps: `rounds` rounds each
 * run every candidate for `steps_per_trial` steps (the scenes advance by
 * all of them); with two or more rounds the first is not counted.  The
 * fastest schedule by wall time is kept for later steps.  Block size
 * (256 particles = 4 wavefronts, one neighbor-list group per wavefront),
 * cells per scan block (1024) and the reorder every step (the reference's
 * bucket order needs it) are fixed by the data layout and not tuned.  No
 * reference counterpart. */
typedef struct sphg_tune_result {
    int32_t density_slack;
    int32_t divergence_slack;
    int32_t replay;                 /* 1: captured segments replayed */
    uint32_t candidates;            /* schedules tried */
    uint32_t steps;                 /* steps run while tuning */
    double ms_per_step;             /* the chosen schedule's mean wall time per step */
    double default_ms_per_step;     /* the default schedule's (1, 0, no replay); -1 if not tried */
} sphg_tune_result;
int sphg_scene_autotune(sphg_scene* scene, uint32_t steps_per_trial, uint32_t rounds, sphg_tune_result* out);
int sphg_batch_autotune(sphg_batch* batch, uint32_t steps_per_trial, uint32_t rounds, sphg_tune_result* out);

/* Device bytes the scene holds once initialised, computed from its fluid and
 * boundary (after set_fluid / set_boundary) before anything is allocated:
 * the admission test of a memory-budgeted serving loop.  No reference
 * counterpart. */
int sphg_scene_plan_bytes(sphg_scene* scene, uint64_t* bytes);

/* Per-kernel device timing with HIP events on the library's stream.
 * enable != 0 starts recording every launch; read returns, for kernel
 * `name`, the launch count and the summed milliseconds since enabling. */
int sphg_ctx_kernel_timing(sphg_ctx* ctx, int32_t enable);
int sphg_ctx_kernel_time(sphg_ctx* ctx, const char* name, uint64_t* launches, double* total_ms);
/* As sphg_ctx_kernel_timing(ctx, 1), recording only the `count` kernels
 * named: each event pair costs the stream a few microseconds, so a
 * throughput measurement times only the kernel it needs. */
int sphg_ctx_kernel_timing_only(sphg_ctx* ctx, const char* const* names, uint32_t count);
/* While timing is on, record only every `stride`-th launch of each kernel
 * (from the first): a sample of the launches, at 1/stride of the event
 * pairs' cost to the stream; launch counts and summed milliseconds then
 * cover the sampled launches.  Enabling timing resets the stride to 1.  The
 * reference's TimeManager counters (src/TimeManager.h:81-94) have no
 * sampling. */
int sphg_ctx_kernel_timing_stride(sphg_ctx* ctx, uint32_t stride);

/* Scene geometry (set-up path, FP64 then rounded to float):
 * AkinciBoundaryModel::sampleCube (src/AkinciBoundayModel.cpp:100-134);
 * returns the count, writes min(count, cap) points. */
int64_t sphg_sample_box(const double* mn, const double* mx, double support_radius, float* out, int64_t cap);
/* AkinciBoundaryModel::sampleSphere (src/AkinciBoundayModel.cpp:136-164). */
int64_t sphg_sample_sphere(const double* center, double radius, double particle_radius, float* out, int64_t cap);
/* PCISPHBoundaryModel::addCube / addSphere (src/PCISPHBoundaryModel.cpp:
 * 7-52, 60-216): points (box surface at ceil(l / 0.45h) intervals; the
 * sphere's points as above) and their normals, inward for an inverted box
 * or sphere; nrm may be NULL.  Returns the point count, or -1 on bad
 * arguments (a box needs a positive extent on every axis). */
int64_t sphg_sample_box_normals(const double* min, const double* max, double support_radius, int32_t inverted,
                                float* xyz, float* nrm, int64_t cap);
int64_t sphg_sample_sphere_normals(const double* center, double radius, double particle_radius, int32_t inverted,
                                   float* xyz, float* nrm, int64_t cap);
/* Simulation::buildFluidBlock (src/Simulation.cpp:512-539) with an optional
 * seeded jitter of +-jitter*2r per coordinate (SplitMix64; 0 = none). */
int64_t sphg_fluid_block(const double* mn, const double* mx, double particle_radius, double jitter, uint64_t seed,
                         float* out, int64_t cap);
/* Lattice sites (block lattice of box [mn, mx]) inside any of the spheres,
 * in lattice order (scene extension; no reference counterpart). */
int64_t sphg_fluid_spheres(const double* mn, const double* mx, double particle_radius, const double* centers,
                           const double* radii, int32_t nspheres, float* out, int64_t cap);

#ifdef __cplusplus
}
#endif

#endif /* SPHG_H */
