// SPH.cpp -- host layer of the reference class surface over libsphg.
#include "SPH.h"

#include <algorithm>
#include <fstream>
#include <cstdio>
#include <sstream>

// ------------------------------------------------------------ helpers

namespace {

void check(int rc, const char* what) {
    if (rc != SPHG_OK) throw SphError(std::string(what) + ": " + sphg_last_error());
}

std::vector<float> toFloat(const std::vector<Vector3r>& v) {
    std::vector<float> out(3 * v.size());
    for (size_t i = 0; i < v.size(); ++i) {
        out[3 * i] = (float)v[i].x;
        out[3 * i + 1] = (float)v[i].y;
        out[3 * i + 2] = (float)v[i].z;
    }
    return out;
}

std::vector<Vector3r> fromFloat(const std::vector<float>& f) {
    std::vector<Vector3r> out(f.size() / 3);
    for (size_t i = 0; i < out.size(); ++i) out[i] = Vector3r(f[3 * i], f[3 * i + 1], f[3 * i + 2]);
    return out;
}

sphg_ctx*& sharedCtx() {
    static sphg_ctx* c = nullptr;
    return c;
}

}  // namespace

// ------------------------------------------------------------- models

void FluidModel::resizeFluid(unsigned int n) {
    mass.resize(n);
    density.resize(n);
    pressure.resize(n);
    r.resize(n);
    v.resize(n);
    a.resize(n);
}

void FluidModel::init(std::vector<Vector3r>& points, std::vector<Vector3r>& velocities) {
    resizeFluid((unsigned int)points.size());
    for (size_t i = 0; i < points.size(); ++i) {
        r[i] = points[i];
        v[i] = i < velocities.size() ? velocities[i] : Vector3r();
    }
    numActiveParticles = (unsigned int)points.size();
}

void FluidModel::setMasses(Real m) { std::fill(mass.begin(), mass.end(), m); }

unsigned int Emitter::SQUARE_EMITTER = SPHG_EMITTER_SQUARE;
unsigned int Emitter::CIRCLE_EMITTER = SPHG_EMITTER_CIRCLE;

sphg_emitter_desc Emitter::desc() const {
    sphg_emitter_desc d{};
    d.type = type;
    d.num_particles = numParticles;
    d.position[0] = r.x; d.position[1] = r.y; d.position[2] = r.z;
    d.velocity = v;
    d.rotation[0] = rot.w; d.rotation[1] = rot.x; d.rotation[2] = rot.y; d.rotation[3] = rot.z;
    d.start_time = startTime;
    d.width = width;
    d.height = height;
    d.spacing = spacing;
    return d;
}

void FluidModel::addEmitter(unsigned int type, unsigned int numParticles, Vector3r pos, Real vel, Quat4r rot,
                            Real startTime, Real w, Real h, Real spacing) {
    emitters.emplace_back(this, type, numParticles, pos, vel, rot, startTime, w, h, spacing);
    resizeFluid(getNumParticles() + numParticles);   // masses: Simulation::init sets every slot
}

void FluidModel::setViscosityForce(Real viscosity, Real boundaryViscosity) {
    Simulation* sim = Simulation::getCurrent();
    const int m = sim->getViscosityMethod();
    ViscosityForce* f = nullptr;
    if (m == Simulation::STANDARD_VISCOSITY_METHOD) f = new StandardViscosityForce(this);
    else if (m == Simulation::ARTIFICIAL_VISCOSITY_METHOD) f = new ArtificialViscosity(this);
    else if (m == Simulation::XSPH_VISCOSITY_METHOD) f = new XSPHViscosity(this);
    if (f) {
        f->init(viscosity, boundaryViscosity);
        npForces.emplace_back(f);
    }
}

void FluidModel::setSurfaceTensionForce(Real c) {
    if (Simulation::getCurrent()->getSurfaceTensionMethod() == Simulation::AKINCI_SURFACE_TENSION_METHOD) {
        auto* f = new SurfaceTensionAkinci(this);
        f->init(c);
        npForces.emplace_back(f);
    }
}

void FluidModel::setAdhesionForce(Real beta) {
    if (Simulation::getCurrent()->getAdhesionMethod() == Simulation::AKINCI_ADHESION_METHOD) {
        auto* f = new AdhesionForce(this);
        f->init(beta);
        npForces.emplace_back(f);
    }
}

void AkinciBoundaryModel::addCube(Vector3r mn, Vector3r mx) {
    const double a[3] = {mn.x, mn.y, mn.z}, b[3] = {mx.x, mx.y, mx.z};
    const double h = Simulation::getCurrent()->getSupportRadius();
    const int64_t n = sphg_sample_box(a, b, h, nullptr, 0);
    if (n < 0) throw SphError("bad box");
    std::vector<float> pts(3 * (size_t)n);
    sphg_sample_box(a, b, h, pts.data(), n);
    auto add = fromFloat(pts);
    const size_t base = r.size();
    resizeBoundary((unsigned int)(base + add.size()));
    std::copy(add.begin(), add.end(), r.begin() + base);
}

void AkinciBoundaryModel::addSphere(Vector3r o, Real radius) {
    const double c[3] = {o.x, o.y, o.z};
    const double pr = Simulation::getCurrent()->getParticleRadius();
    const int64_t n = sphg_sample_sphere(c, radius, pr, nullptr, 0);
    if (n < 0) throw SphError("bad sphere");
    std::vector<float> pts(3 * (size_t)n);
    sphg_sample_sphere(c, radius, pr, pts.data(), n);
    auto add = fromFloat(pts);
    const size_t base = r.size();
    resizeBoundary((unsigned int)(base + add.size()));
    std::copy(add.begin(), add.end(), r.begin() + base);
}

void PCISPHBoundaryModel::addCube(Vector3r mn, Vector3r mx, bool inverted) {
    const double a[3] = {mn.x, mn.y, mn.z}, b[3] = {mx.x, mx.y, mx.z};
    const double h = Simulation::getCurrent()->getSupportRadius();
    const int64_t k = sphg_sample_box_normals(a, b, h, inverted ? 1 : 0, nullptr, nullptr, 0);
    if (k < 0) throw SphError("bad box (the PCISPH boundary sampler needs a positive extent on every axis)");
    std::vector<float> pts(3 * (size_t)k), nrm(3 * (size_t)k);
    sphg_sample_box_normals(a, b, h, inverted ? 1 : 0, pts.data(), nrm.data(), k);
    auto p = fromFloat(pts), q = fromFloat(nrm);
    const size_t base = r.size();
    resizeBoundary((unsigned int)(base + p.size()));
    std::copy(p.begin(), p.end(), r.begin() + base);
    std::copy(q.begin(), q.end(), n.begin() + base);
}

void PCISPHBoundaryModel::addSphere(Vector3r o, Real radius, bool inverted) {
    const double c[3] = {o.x, o.y, o.z};
    const double pr = Simulation::getCurrent()->getParticleRadius();
    const int64_t k = sphg_sample_sphere_normals(c, radius, pr, inverted ? 1 : 0, nullptr, nullptr, 0);
    if (k < 0) throw SphError("bad sphere");
    std::vector<float> pts(3 * (size_t)k), nrm(3 * (size_t)k);
    sphg_sample_sphere_normals(c, radius, pr, inverted ? 1 : 0, pts.data(), nrm.data(), k);
    auto p = fromFloat(pts), q = fromFloat(nrm);
    const size_t base = r.size();
    resizeBoundary((unsigned int)(base + p.size()));
    std::copy(p.begin(), p.end(), r.begin() + base);
    std::copy(q.begin(), q.end(), n.begin() + base);
}

// ------------------------------------------------------------ solvers

GpuSolver::~GpuSolver() {
    try {
        flushFrames();
    } catch (const std::exception&) {
    }
    for (auto& f : frames)
        if (f.snap) sphg_snapshot_destroy(f.snap);
    if (scene) sphg_scene_destroy(scene);
}

void GpuSolver::queueFrame(const std::string& filename) {
    if (!scene) throw SphError("frame before init");
    FrameSlot& f = frames[nextFrame];
    nextFrame ^= 1;
    if (f.writer.joinable()) f.writer.join();
    if (!f.snap) check(sphg_snapshot_create(scene, &f.snap), "sphg_snapshot_create");
    check(sphg_snapshot_begin(f.snap), "sphg_snapshot_begin");
    uint32_t cap = 0;
    check(sphg_scene_capacity(scene, &cap), "capacity");
    sphg_snapshot* snap = f.snap;
    f.writer = std::thread([this, snap, filename, cap]() {
        try {
            const float *x = nullptr, *v = nullptr;
            uint32_t n = 0;
            double t = 0, dt = 0;
            if (sphg_snapshot_wait(snap, &x, &v, &n, &t, &dt)) throw SphError(sphg_last_error());
            std::ofstream out(filename);
            if (!out.is_open()) throw SphError("cannot write " + filename);
            out << SceneLoader::frameJsonFrom(cap, n, t, dt, x, v);
            if (!out) throw SphError("writing " + filename + " failed");
        } catch (const std::exception& e) {
            std::lock_guard<std::mutex> g(frameMutex);
            if (frameError.empty()) frameError = e.what();
        }
    });
}

void GpuSolver::flushFrames() {
    for (auto& f : frames)
        if (f.writer.joinable()) f.writer.join();
    std::lock_guard<std::mutex> g(frameMutex);
    if (!frameError.empty()) {
        std::string e;
        std::swap(e, frameError);
        throw SphError("frame output: " + e);
    }
}

void GpuSolver::fillDesc(sphg_scene_desc& d) const {
    Simulation* sim = Simulation::getCurrent();
    d.method = method;
    d.boundary_method = sim->getBoundaryHandlingMethod();
    // the cube method's box: one CubeBoundaryModel (the reference's scene
    // loader creates none, src/Simulation.cpp:435-486, so a scene file with
    // boundaryMethod 0 has no boundary at all)
    if (d.boundary_method == Simulation::CUBE_BOUNDARY_METHOD) {
        d.cube_normal_fct = 0.0;
        d.cube_tangential_fct = 1.0;
        unsigned int cubes = 0;
        for (unsigned int b = 0; b < sim->numberBoundaryModels(); ++b)
            if (auto* c = dynamic_cast<CubeBoundaryModel*>(sim->getBoundaryModel(b))) {
                if (++cubes > 1) throw SphError("the GPU solver takes one CubeBoundaryModel");
                d.cube_enabled = 1;
                const Vector3r mn = c->getMin(), mx = c->getMax();
                d.cube_min[0] = mn.x; d.cube_min[1] = mn.y; d.cube_min[2] = mn.z;
                d.cube_max[0] = mx.x; d.cube_max[1] = mx.y; d.cube_max[2] = mx.z;
                d.cube_normal_fct = c->getNormalFct();
                d.cube_tangential_fct = c->getTangentialFct();
            }
    }
    // the PCISPH method's velocity factors: one PCISPHBoundaryModel (its
    // corrections run model after model in the reference,
    // src/PCISPHSolver.cpp:67-74, which one model on the device reproduces)
    d.pb_normal_fct = 0.0;
    d.pb_tangential_fct = 1.0;
    if (d.boundary_method == Simulation::PCISPH_BOUNDARY_METHOD) {
        unsigned int models = 0;
        for (unsigned int b = 0; b < sim->numberBoundaryModels(); ++b)
            if (auto* m = dynamic_cast<PCISPHBoundaryModel*>(sim->getBoundaryModel(b))) {
                if (++models > 1) throw SphError("the GPU solver takes one PCISPHBoundaryModel (one Boundary entry)");
                d.pb_normal_fct = m->getNormalFct();
                d.pb_tangential_fct = m->getTangentialFct();
            }
    }
    d.particle_radius = sim->getParticleRadius();
    d.density0 = sim->numberFluidModels() ? sim->getFluidModel(0)->getRefDensity() : 1000.0;
    const Vector3r g = sim->getGravity();
    d.gravity[0] = g.x;
    d.gravity[1] = g.y;
    d.gravity[2] = g.z;
    d.time_step = sim->getTimeStep();
    d.min_time_step = sim->getMinTimeStep();
    d.max_time_step = sim->getMaxTimeStep();
    d.fps = sim->getFPS();
    d.eta = maxError;
    d.min_iterations = minIterations;
    d.max_iterations = maxIterations;
    d.viscosity_method = -1;
    d.surface_tension_method = -1;
    d.adhesion_method = -1;
    d.max_neighbors = sim->getMaxNeighbors();
    if (sim->numberFluidModels()) {
        FluidModel* fm = sim->getFluidModel(0);
        for (unsigned int i = 0; i < fm->numberNonPressureForces(); ++i) {
            NonPressureForce* f = fm->getNonPressureForce(i);
            if (auto* x = dynamic_cast<ViscosityForce*>(f)) {
                d.viscosity_method = x->method();
                d.viscosity = x->getViscosity();
                d.boundary_viscosity = x->getBoundaryViscosity();
            } else if (auto* s = dynamic_cast<SurfaceTensionAkinci*>(f)) {
                d.surface_tension_method = SPHG_SURFACE_TENSION_AKINCI;
                d.surface_tension = s->getSurfaceTension();
            } else if (auto* a = dynamic_cast<AdhesionForce*>(f)) {
                d.adhesion_method = SPHG_ADHESION_AKINCI;
                d.adhesion = a->getAdhesion();
            }
        }
    }
}

void WCSPHSolver::fillDesc(sphg_scene_desc& d) const {
    GpuSolver::fillDesc(d);
    d.stiffness = stiffness;
    d.gamma = gamma;
}

void DFSPHSolver::fillDesc(sphg_scene_desc& d) const {
    GpuSolver::fillDesc(d);
    d.cfl_factor = cflFactor;
    d.eta_v = maxErrorV;
    d.min_iterations_v = minIterationsV;
    d.max_iterations_v = maxIterationsV;
}

// Simulation::init -> solver->init(): one fluid model and one Akinci
// boundary model are uploaded (the reference's own limits, its model-id
// handling at src/Simulation.h:24-29 and src/AkinciBoundayModel.cpp:32).
void GpuSolver::init() {
    Simulation* sim = Simulation::getCurrent();
    const int bm = sim->getBoundaryHandlingMethod();
    if (bm != Simulation::AKINCI_BOUNDARY_METHOD && bm != Simulation::CUBE_BOUNDARY_METHOD &&
        bm != Simulation::PCISPH_BOUNDARY_METHOD)
        throw SphError("the GPU solver needs boundaryMethod 0 (cube), 1 (PCISPH) or 2 (Akinci)");
    if (sim->numberFluidModels() != 1) throw SphError("the GPU solver runs one fluid model per scene");
    sphg_ctx*& c = sharedCtx();
    if (!c) check(sphg_ctx_create(sim->getDevice(), &c), "sphg_ctx_create");
    ctx = c;
    sphg_scene_desc d{};
    fillDesc(d);
    check(sphg_scene_create(ctx, &d, &scene), "sphg_scene_create");
    FluidModel* fm = sim->getFluidModel(0);
    const unsigned int n0 = fm->getNumActiveParticles();
    std::vector<float> x = toFloat(fm->getPositions()), v = toFloat(fm->getVelocitiesVec());
    if (n0) check(sphg_scene_set_fluid(scene, x.data(), v.data(), n0), "set_fluid");
    for (unsigned int e = 0; e < fm->numberEmitters(); ++e) {
        const sphg_emitter_desc ed = fm->getEmitter(e).desc();
        check(sphg_scene_add_emitter(scene, &ed), "add_emitter");
    }
    std::vector<Vector3r> bpts, bnrm;
    if (bm == Simulation::AKINCI_BOUNDARY_METHOD || bm == Simulation::PCISPH_BOUNDARY_METHOD)
        for (unsigned int b = 0; b < sim->numberBoundaryModels(); ++b) {
            auto& p = sim->getBoundaryModel(b)->getPosition();
            bpts.insert(bpts.end(), p.begin(), p.end());
            if (auto* m = dynamic_cast<PCISPHBoundaryModel*>(sim->getBoundaryModel(b)))
                bnrm.insert(bnrm.end(), m->getNormals().begin(), m->getNormals().end());
        }
    std::vector<float> bx = toFloat(bpts);
    check(sphg_scene_set_boundary(scene, bx.data(), (uint32_t)bpts.size()), "set_boundary");
    if (bm == Simulation::PCISPH_BOUNDARY_METHOD) {
        std::vector<float> bn = toFloat(bnrm);
        check(sphg_scene_set_boundary_normals(scene, bn.data(), (uint32_t)bnrm.size()), "set_boundary_normals");
    }
    check(sphg_scene_init(scene), "sphg_scene_init");
    // boundary volumes back into the Akinci models
    std::vector<float> vol(std::max<size_t>(bpts.size(), 1));
    if (bm != Simulation::PCISPH_BOUNDARY_METHOD) check(sphg_scene_get(scene, SPHG_FIELD_BVOL, vol.data()), "get volumes");
    size_t off = 0;
    for (unsigned int b = 0; bm == Simulation::AKINCI_BOUNDARY_METHOD && b < sim->numberBoundaryModels(); ++b) {
        auto* abm = dynamic_cast<AkinciBoundaryModel*>(sim->getBoundaryModel(b));
        const unsigned int nb = sim->getBoundaryModel(b)->getNumParticles();
        if (abm)
            for (unsigned int i = 0; i < nb; ++i) abm->getVolume(i) = vol[off + i];
        off += nb;
    }
    check(sphg_scene_get_scalars(scene, &last), "scalars");
    syncToHost();
}

void GpuSolver::step() { stepMany(1); }

void GpuSolver::stepMany(unsigned int k) {
    if (!scene) throw SphError("step before init");
    sphg_step_stats st{};
    check(sphg_scene_step(scene, k, &st), "sphg_scene_step");
    gpuMs += st.gpu_ms;
    check(sphg_scene_get_scalars(scene, &last), "scalars");
    steps = last.steps;
    iterations = (unsigned int)last.iterations;
    sumIterations = last.sum_iterations;
    maxVel = last.max_vel;
    Simulation* sim = Simulation::getCurrent();
    sim->setTimeStep(last.time_step);
    sim->setTime(last.time);
    if (sim->getFluidModel(0)->numberEmitters()) {
        uint32_t n = 0;
        check(sphg_scene_info(scene, &n, nullptr, nullptr, nullptr), "info");
        sim->getFluidModel(0)->setNumActiveParticles(n);
    }
    hostFresh = false;
}

void GpuSolver::syncToHost() {
    if (!scene || hostFresh) return;
    Simulation* sim = Simulation::getCurrent();
    FluidModel* fm = sim->getFluidModel(0);
    // the active particles; slots not yet emitted keep their zero state
    uint32_t n = 0;
    check(sphg_scene_info(scene, &n, nullptr, nullptr, nullptr), "info");
    fm->setNumActiveParticles(n);
    std::vector<float> buf(3 * (size_t)n);
    auto put = [&](std::vector<Vector3r>& dst) {
        for (uint32_t i = 0; i < n; ++i) dst[i] = Vector3r(buf[3 * i], buf[3 * i + 1], buf[3 * i + 2]);
    };
    check(sphg_scene_get(scene, SPHG_FIELD_X, buf.data()), "get x");
    put(fm->getPositions());
    check(sphg_scene_get(scene, SPHG_FIELD_V, buf.data()), "get v");
    put(fm->getVelocitiesVec());
    std::vector<float> s(n);
    check(sphg_scene_get(scene, SPHG_FIELD_RHO, s.data()), "get rho");
    for (unsigned int i = 0; i < n; ++i) fm->getDensity(i) = s[i];
    if (method != SPHG_DFSPH) {
        check(sphg_scene_get(scene, SPHG_FIELD_P, s.data()), "get p");
        for (unsigned int i = 0; i < n; ++i) fm->getPressure(i) = s[i];
    }
    // PCISPH boundary method: the boundary densities and pressures the
    // WCSPH and PCISPH steps form (DFSPH does not use them)
    if (sim->getBoundaryHandlingMethod() == Simulation::PCISPH_BOUNDARY_METHOD && method != SPHG_DFSPH)
        for (unsigned int b = 0; b < sim->numberBoundaryModels(); ++b)
            if (auto* m = dynamic_cast<PCISPHBoundaryModel*>(sim->getBoundaryModel(b))) {
                const unsigned int nb = m->getNumParticles();
                std::vector<float> bd(std::max(nb, 1u)), bp(std::max(nb, 1u));
                check(sphg_scene_get(scene, SPHG_FIELD_BDENSITY, bd.data()), "get boundary densities");
                check(sphg_scene_get(scene, SPHG_FIELD_BPRESSURE, bp.data()), "get boundary pressures");
                for (unsigned int i = 0; i < nb; ++i) {
                    m->getDensity(i) = bd[i];
                    m->getPressure(i) = bp[i];
                }
            }
    hostFresh = true;
}

// --------------------------------------------------------- simulation

Simulation* Simulation::current = nullptr;
const int Simulation::WCSPH_METHOD = 0;
const int Simulation::PCISPH_METHOD = 1;
const int Simulation::DFSPH_METHOD = 2;
const int Simulation::CUBE_BOUNDARY_METHOD = 0;
const int Simulation::PCISPH_BOUNDARY_METHOD = 1;
const int Simulation::AKINCI_BOUNDARY_METHOD = 2;
const int Simulation::STANDARD_VISCOSITY_METHOD = 0;
const int Simulation::ARTIFICIAL_VISCOSITY_METHOD = 1;
const int Simulation::XSPH_VISCOSITY_METHOD = 2;
const int Simulation::AKINCI_SURFACE_TENSION_METHOD = 3;
const int Simulation::AKINCI_ADHESION_METHOD = 4;

Simulation* Simulation::getCurrent() {
    if (!current) current = new Simulation();
    return current;
}
void Simulation::setCurrent(Simulation* s) { current = s; }
bool Simulation::hasCurrent() { return current != nullptr; }
void Simulation::destroyCurrent() {
    delete current;
    current = nullptr;
}

void Simulation::setSimulationMethod(int method) {
    if (method == current_method || method < 0) return;
    if (method > 2) throw SphError("simulationMethod must be 0 (WCSPH), 1 (PCISPH) or 2 (DFSPH)");
    current_method = method;
    if (method == WCSPH_METHOD) solver.reset(new WCSPHSolver());
    else if (method == PCISPH_METHOD) solver.reset(new PCISPHSolver());
    else solver.reset(new DFSPHSolver());
}

FluidModel* Simulation::addFluidModel(std::vector<Vector3r>& p, std::vector<Vector3r>& v) {
    fluidModels.emplace_back(new FluidModel());
    fluidModels.back()->init(p, v);
    return fluidModels.back().get();
}
FluidModel* Simulation::addFluidModel() {
    fluidModels.emplace_back(new FluidModel());
    return fluidModels.back().get();
}

void Simulation::initMasses() {
    const Real d = 2.0 * particleRadius;
    const Real vol = std::pow(d, 3.0);
    for (auto& fm : fluidModels) fm->setMasses(fm->getRefDensity() * vol);
}

void Simulation::computeNonPressureForces() {
    for (auto& fm : fluidModels)
        for (unsigned int i = 0; i < fm->numberNonPressureForces(); ++i) fm->getNonPressureForce(i)->step();
}

void Simulation::init() {
    if (!solver) throw SphError("no simulation method set");
    initMasses();
    solver->init();
}

bool Simulation::step() {
    const Real time = getTime();
    bool save = false;
    if (time >= getStartTime() && time <= getEndTime()) {
        solver->step();
        save = tm.hasToSave();
        if (save) {
            if (activeSave) SceneLoader::writeFluid();
            tm.setFrame(tm.getFrame() + 1);
        }
    }
    return save;
}

void Simulation::run(unsigned int maxSteps) {
    unsigned int k = 0;
    auto* gs = dynamic_cast<GpuSolver*>(solver.get());
    while (getTime() >= getStartTime() && getTime() <= getEndTime()) {
        // Steps that end at least one step before the next frame time and
        // the end time leave the per-step checks of step() nothing to do, and
        // go to the device in one call.  Their time steps are bounded from
        // the current one: WCSPH's and PCISPH's is fixed, and DFSPH's CFL
        // rule (src/DFSPHSolver.cpp:827-870) grows it by at most 1.001 per
        // step before clamping it to [minTimeStep, maxTimeStep].
        unsigned int run = 1;
        if (gs) {
            const unsigned int cap = maxSteps ? maxSteps - k : 1000000u;
            const bool cfl = gs->getMethod() == SPHG_DFSPH;
            const Real tEnd = std::min(tm.nextSaveTime(), getEndTime());
            Real t = getTime(), dt = gs->scalars().time_step;
            unsigned int fit = 0;
            while (fit < cap && dt > 0) {
                const Real next = cfl ? std::min(getMaxTimeStep(), std::max(getMinTimeStep(), dt * 1.0011)) : dt;
                if (!(next > 0) || !(t + 2 * next < tEnd)) break;
                t += next;
                dt = next;
                ++fit;
            }
            if (fit >= 2) run = fit;
        }
        if (run > 1) gs->stepMany(run);
        else step();
        k += run;
        if (maxSteps && k >= maxSteps) break;
    }
    if (auto* gs = dynamic_cast<GpuSolver*>(solver.get())) gs->flushFrames();
}

FluidModel* Simulation::buildFluidBlock(const std::vector<BlockInfo>& blocks, const std::vector<SphereInfo>& spheres,
                                        Vector3r smin, Vector3r smax) {
    std::vector<Vector3r> pos, vel;
    for (const BlockInfo& b : blocks) {
        const double mn[3] = {b.min.x, b.min.y, b.min.z}, mx[3] = {b.max.x, b.max.y, b.max.z};
        const int64_t n = sphg_fluid_block(mn, mx, particleRadius, b.jitter, b.seed, nullptr, 0);
        if (n < 0) throw SphError("bad fluid block");
        std::vector<float> p(3 * (size_t)std::max<int64_t>(n, 1));
        sphg_fluid_block(mn, mx, particleRadius, b.jitter, b.seed, p.data(), n);
        p.resize(3 * (size_t)n);
        auto add = fromFloat(p);
        pos.insert(pos.end(), add.begin(), add.end());
    }
    if (!spheres.empty()) {
        std::vector<double> c, r;
        for (const SphereInfo& s : spheres) {
            c.insert(c.end(), {s.center.x, s.center.y, s.center.z});
            r.push_back(s.radius);
        }
        const double mn[3] = {smin.x, smin.y, smin.z}, mx[3] = {smax.x, smax.y, smax.z};
        const int64_t n = sphg_fluid_spheres(mn, mx, particleRadius, c.data(), r.data(), (int)r.size(), nullptr, 0);
        if (n < 0) throw SphError("bad fluid spheres");
        std::vector<float> p(3 * (size_t)std::max<int64_t>(n, 1));
        sphg_fluid_spheres(mn, mx, particleRadius, c.data(), r.data(), (int)r.size(), p.data(), n);
        p.resize(3 * (size_t)n);
        auto add = fromFloat(p);
        pos.insert(pos.end(), add.begin(), add.end());
    }
    vel.assign(pos.size(), Vector3r());
    return addFluidModel(pos, vel);
}

std::string Simulation::describe() {
    static const char* solvers[] = {"Weakly Compressible SPH", "Predictive-Corrective Incompressible SPH",
                                    "Divergence-Free SPH"};
    std::ostringstream o;
    o << "Pressure solver:            " << (current_method >= 0 ? solvers[current_method] : "none") << " (GPU)\n";
    o << "Fluid models:               " << fluidModels.size() << "\n";
    for (size_t i = 0; i < fluidModels.size(); ++i)
        o << "  particles in fm " << i + 1 << ":        " << fluidModels[i]->getNumParticles() << "\n";
    o << "Boundary models:            " << boundaryModels.size() << "\n";
    for (size_t i = 0; i < boundaryModels.size(); ++i)
        o << "  particles in bm " << i + 1 << ":        " << boundaryModels[i]->getNumParticles() << "\n";
    o << "Particle radius:            " << particleRadius << " m\n";
    o << "Support radius:             " << supportRadius << " m\n";
    o << "Gravity:                    (" << gravity.x << "," << gravity.y << "," << gravity.z << ") m/s^2\n";
    return o.str();
}
