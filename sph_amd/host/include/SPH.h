// SPH.h -- the host-side class surface of the reference SPH solver
// (josemanuel28v/SPH), backed by the gfx950 solver through the C ABI of
// libsphg (include/sphg.h).
//
// Classes and methods keep the reference's names and meaning so scene code
// written against it reads the same:
//   Simulation        src/Simulation.h:148-317   (singleton, method enums)
//   SPHSolver         src/SPHSolver.h:6-50       (abstract solver)
//   WCSPHSolver       src/WCSPHSolver.h:8-34
//   PCISPHSolver      src/PCISPHSolver.h:8-49
//   DFSPHSolver       src/DFSPHSolver.h:8-70
//   FluidModel        src/FluidModel.h:10-82
//   BoundaryModel     src/BoundaryModel.h:7-30
//   AkinciBoundaryModel src/AkinciBoundaryModel.h:8-30
//   NonPressureForce  src/NonPressureForce.h:6-19 (+ XSPHViscosity,
//                     SurfaceTensionAkinci, AdhesionForce)
//   TimeManager       src/TimeManager.h:15-117
// What changes: the solver's step runs on the device; FluidModel's arrays
// are a host mirror refreshed from the device at defined points (after
// init, before a frame is written, on syncToHost()); force objects carry
// coefficients that the device kernels apply (their step() does nothing on
// the host).  Errors are reported as SphError exceptions inside the host
// layer (the C ABI below never throws).
#pragma once

#include <chrono>
#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <thread>
#include <string>
#include <vector>

#include "../../../include/sphg.h"

typedef double Real;

struct Vector3r {
    Real x = 0, y = 0, z = 0;
    Vector3r() = default;
    Vector3r(Real a, Real b, Real c) : x(a), y(b), z(c) {}
    explicit Vector3r(Real s) : x(s), y(s), z(s) {}
    Vector3r operator+(const Vector3r& o) const { return {x + o.x, y + o.y, z + o.z}; }
    Vector3r operator-(const Vector3r& o) const { return {x - o.x, y - o.y, z - o.z}; }
    Vector3r operator*(Real s) const { return {x * s, y * s, z * s}; }
};

// glm::quat's (w, x, y, z) constructor order (src/SceneLoader.cpp:131).
struct Quat4r {
    Real w = 1, x = 0, y = 0, z = 0;
    Quat4r() = default;
    Quat4r(Real a, Real b, Real c, Real d) : w(a), x(b), y(c), z(d) {}
};

class SphError : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};

class FluidModel;
class Simulation;

// ------------------------------------------------------------ forces

class NonPressureForce {
protected:
    FluidModel* fm;

public:
    explicit NonPressureForce(FluidModel* f) : fm(f) {}
    virtual ~NonPressureForce() {}
    // The device applies the force inside the solver step; nothing to do here.
    virtual void step() {}
    virtual void resize(const unsigned int) {}
};

// The three viscosity forces (src/FluidModel.cpp:73-96) carry their two
// coefficients; the device kernel applies them.
class ViscosityForce : public NonPressureForce {
protected:
    Real viscosity = 0, boundaryViscosity = 0;

public:
    using NonPressureForce::NonPressureForce;
    void init(Real v, Real bv) { viscosity = v; boundaryViscosity = bv; }
    Real getViscosity() const { return viscosity; }
    Real getBoundaryViscosity() const { return boundaryViscosity; }
    virtual int method() const = 0;   // SPHG_VISCOSITY_*
};
class StandardViscosityForce : public ViscosityForce {   // src/StandardViscosityForce.h:12-84
public:
    using ViscosityForce::ViscosityForce;
    int method() const override { return SPHG_VISCOSITY_STANDARD; }
};
class ArtificialViscosity : public ViscosityForce {   // src/ArtificialViscosity.h:11-90
public:
    using ViscosityForce::ViscosityForce;
    int method() const override { return SPHG_VISCOSITY_ARTIFICIAL; }
};
class XSPHViscosity : public ViscosityForce {   // src/XSPHViscosity.h:13-74
public:
    using ViscosityForce::ViscosityForce;
    int method() const override { return SPHG_VISCOSITY_XSPH; }
};

class SurfaceTensionAkinci : public NonPressureForce {   // src/SurfaceTensionAkinci.h:15-104
protected:
    Real stCoef = 0;

public:
    using NonPressureForce::NonPressureForce;
    void init(Real c) { stCoef = c; }
    Real getSurfaceTension() const { return stCoef; }
};

class AdhesionForce : public NonPressureForce {   // src/AdhesionForce.h:14-66
protected:
    Real beta = 0;

public:
    using NonPressureForce::NonPressureForce;
    void init(Real b) { beta = b; }
    Real getAdhesion() const { return beta; }
};

// ------------------------------------------------------------ models

// A particle emitter (src/Emitter.h:9-55).  It keeps the reference's
// parameters; the device activates its groups inside the solver step
// (sphg_scene_add_emitter), so emitParticles() does nothing on the host.
class Emitter {
protected:
    unsigned int type, numParticles;
    Vector3r r;
    Real v;
    Quat4r rot;
    Real startTime, width, height, spacing;

public:
    static unsigned int SQUARE_EMITTER;   // src/Emitter.cpp:7
    static unsigned int CIRCLE_EMITTER;   // src/Emitter.cpp:8
    Emitter(FluidModel*, unsigned int type, unsigned int numParticles, Vector3r r, Real v, Quat4r rot, Real startTime,
            Real w, Real h, Real spacing)
        : type(type), numParticles(numParticles), r(r), v(v), rot(rot), startTime(startTime), width(w), height(h),
          spacing(spacing) {}
    void emitParticles() {}
    unsigned int getType() const { return type; }
    unsigned int getNumParticles() const { return numParticles; }
    sphg_emitter_desc desc() const;
};

class FluidModel {   // src/FluidModel.h:10-82
protected:
    Real density_0 = 1000;
    std::vector<Real> mass, density, pressure;
    std::vector<Vector3r> r, v, a;
    unsigned int numActiveParticles = 0;
    std::vector<std::unique_ptr<NonPressureForce>> npForces;
    std::vector<Emitter> emitters;

public:
    void init(std::vector<Vector3r>& points, std::vector<Vector3r>& velocities);
    void resizeFluid(unsigned int n);
    void setMasses(Real m);
    void setRefDensity(Real d) { density_0 = d; }
    Real& getRefDensity() { return density_0; }
    Real& getMass(unsigned int i) { return mass[i]; }
    Real& getDensity(unsigned int i) { return density[i]; }
    Real& getPressure(unsigned int i) { return pressure[i]; }
    Vector3r& getPosition(unsigned int i) { return r[i]; }
    Vector3r& getVelocity(unsigned int i) { return v[i]; }
    Vector3r& getAcceleration(unsigned int i) { return a[i]; }
    std::vector<Vector3r>& getPositions() { return r; }
    std::vector<Vector3r>& getVelocitiesVec() { return v; }
    std::vector<Real>& getDensities() { return density; }
    std::vector<Real>& getPressures() { return pressure; }
    unsigned int getNumParticles() const { return (unsigned int)mass.size(); }
    unsigned int getNumActiveParticles() const { return numActiveParticles; }
    void setNumActiveParticles(unsigned int n) { if (n <= getNumParticles()) numActiveParticles = n; }   // src/FluidModel.h:63

    // src/FluidModel.cpp:124-143: the capacity grows by the emitter's budget
    void addEmitter(unsigned int type, unsigned int numParticles, Vector3r r, Real v, Quat4r rot, Real startTime,
                    Real w, Real h, Real spacing);
    unsigned int numberEmitters() const { return (unsigned int)emitters.size(); }
    Emitter& getEmitter(unsigned int i) { return emitters[i]; }

    void setViscosityForce(Real viscosity, Real boundaryViscosity);   // src/FluidModel.cpp:73-100
    void setSurfaceTensionForce(Real stCoef);                           // :102-111
    void setAdhesionForce(Real beta);                                   // :113-122
    NonPressureForce* getNonPressureForce(unsigned int i) { return npForces[i].get(); }
    unsigned int numberNonPressureForces() const { return (unsigned int)npForces.size(); }
};

class BoundaryModel {   // src/BoundaryModel.h:7-30
protected:
    std::vector<Vector3r> r;

public:
    virtual ~BoundaryModel() {}
    virtual void init(std::vector<Vector3r>& points) { r = points; }
    virtual void resizeBoundary(unsigned int n) { r.resize(n); }
    unsigned int getNumParticles() const { return (unsigned int)r.size(); }
    unsigned int size() const { return (unsigned int)r.size(); }
    Vector3r& getPosition(unsigned int i) { return r[i]; }
    std::vector<Vector3r>& getPosition() { return r; }
};

class AkinciBoundaryModel : public BoundaryModel {   // src/AkinciBoundaryModel.h:8-30
protected:
    std::vector<Real> volume;

public:
    void resizeBoundary(unsigned int n) override { BoundaryModel::resizeBoundary(n); volume.resize(n); }
    Real& getVolume(unsigned int i) { return volume[i]; }
    std::vector<Real>& getVolumes() { return volume; }
    void addCube(Vector3r min, Vector3r max);         // src/AkinciBoundayModel.cpp:42-58
    void addSphere(Vector3r origin, Real radius);     // src/AkinciBoundayModel.cpp:60-76
};

// The PCISPH boundary method's model (src/PCISPHBoundaryModel.h:10-52):
// points sampled at 0.45h on boxes and spheres with their normals (inverted
// keeps the sampler's inward normals), boundary mass and rest density, the
// normal and tangential velocity factors of the collision projection, and
// a host mirror of the boundary densities and pressures.  The device
// computes those and runs the projection inside the solver step.
class PCISPHBoundaryModel : public BoundaryModel {
protected:
    std::vector<Vector3r> n;
    std::vector<Real> density, pressure;
    Real density0 = 1000, mass = 0, normalFct = 0, tangentialFct = 1;

public:
    void setNormalFct(Real f) { normalFct = f; }
    void setTangentialFct(Real f) { tangentialFct = f; }
    void setRefDensity(Real d) { density0 = d; }
    void setMass(Real m) { mass = m; }
    Real getNormalFct() const { return normalFct; }
    Real getTangentialFct() const { return tangentialFct; }
    Real getRefDensity() const { return density0; }
    Real getMass() const { return mass; }
    Real& getDensity(unsigned int i) { return density[i]; }
    Real& getPressure(unsigned int i) { return pressure[i]; }
    Vector3r& getNormal(unsigned int i) { return n[i]; }
    std::vector<Vector3r>& getNormals() { return n; }
    void resizeBoundary(unsigned int k) override {
        BoundaryModel::resizeBoundary(k);
        n.resize(k);
        density.resize(k);
        pressure.resize(k);
    }
    void addCube(Vector3r min, Vector3r max, bool inverted);         // src/PCISPHBoundaryModel.cpp:7-27
    void addSphere(Vector3r origin, Real radius, bool inverted);     // src/PCISPHBoundaryModel.cpp:29-52
};

// The box of the cube boundary method: points [min, max] (init), normal and
// tangential velocity factors 0 and 1 by default.  The clamp runs on the
// device after each step's integration.
class CubeBoundaryModel : public BoundaryModel {   // src/CubeBoundaryModel.h:7-32
protected:
    Real normalFct = 0, tangentialFct = 1;

public:
    void init(std::vector<Vector3r>& points) override {   // src/CubeBoundaryModel.cpp:5-14
        BoundaryModel::init(points);
        normalFct = 0.0;
        tangentialFct = 1.0;
    }
    Vector3r& getMin() { return r.at(0); }
    Vector3r& getMax() { return r.at(1); }
    void setNormalFct(Real f) { normalFct = f; }
    void setTangentialFct(Real f) { tangentialFct = f; }
    Real getNormalFct() const { return normalFct; }
    Real getTangentialFct() const { return tangentialFct; }
};

// ------------------------------------------------------------ solvers

class SPHSolver {   // src/SPHSolver.h:6-50
protected:
    unsigned int iterations = 0;
    unsigned int minIterations = 3;
    unsigned int maxIterations = 100;
    unsigned int steps = 0;
    unsigned int sumIterations = 0;
    Real maxError = 0.01;
    Real maxVel = 0;

public:
    virtual ~SPHSolver() {}
    virtual void init() = 0;
    virtual void step() = 0;
    virtual void resizeData() = 0;
    // Phases of the reference's base class; on the device they are part of
    // step() and are not callable one by one.
    virtual void computeDensities() {}
    virtual void integrate() {}
    virtual void insertFluidParticles() {}
    virtual void insertBoundaryParticles() {}
    virtual void neighborhoodSearch() {}

    void setMaxError(Real e) { maxError = e; }
    void setMinIterations(unsigned int it) { minIterations = it; }
    void setMaxIterations(unsigned int it) { maxIterations = it; }
    Real getMaxError() const { return maxError; }
    Real getMaxVel() const { return maxVel; }
    unsigned int getSteps() const { return steps; }
    unsigned int getMinIterations() const { return minIterations; }
    unsigned int getMaxIterations() const { return maxIterations; }
    unsigned int getIterations() const { return iterations; }
};

// Device-backed solver: one sphg_scene per Simulation.
class GpuSolver : public SPHSolver {
protected:
    int method;
    sphg_ctx* ctx = nullptr;
    sphg_scene* scene = nullptr;
    bool hostFresh = false;
    sphg_scene_scalars last{};
    double gpuMs = 0;
    // frames in flight: a device snapshot each, written by its own thread
    struct FrameSlot {
        sphg_snapshot* snap = nullptr;
        std::thread writer;
    };
    FrameSlot frames[2];
    int nextFrame = 0;
    std::mutex frameMutex;
    std::string frameError;

public:
    explicit GpuSolver(int method_) : method(method_) {}
    ~GpuSolver() override;
    void init() override;
    void step() override;
    // k steps in one library call, leaving the host state as k calls of
    // step() do (the caller knows that no frame falls due among them)
    void stepMany(unsigned int k);
    void resizeData() override {}
    // Refresh FluidModel's host mirror (positions, velocities, densities,
    // pressures) from the device, in particle-id order.
    void syncToHost();
    // A frame of the current state (SceneLoader::writeFluid's content) to
    // `filename` without stopping the host: the read-back runs beside the
    // next steps and a writer thread formats and writes the file; two
    // frames are in flight at most.
    void queueFrame(const std::string& filename);
    // Waits for every queued frame; a writer's error is rethrown here.
    void flushFrames();
    const sphg_scene_scalars& scalars() const { return last; }
    double deviceMilliseconds() const { return gpuMs; }
    sphg_scene* handle() { return scene; }
    virtual void fillDesc(sphg_scene_desc& d) const;
    int getMethod() const { return method; }
};

class WCSPHSolver : public GpuSolver {   // src/WCSPHSolver.h:8-34
protected:
    Real stiffness = 1000, gamma = 7;

public:
    WCSPHSolver() : GpuSolver(SPHG_WCSPH) {}
    void setStiffness(Real s) { stiffness = s; }
    void setGamma(Real g) { gamma = g; }
    Real getStiffness() const { return stiffness; }
    Real getGamma() const { return gamma; }
    void fillDesc(sphg_scene_desc& d) const override;
};

class PCISPHSolver : public GpuSolver {   // src/PCISPHSolver.h:8-49
public:
    PCISPHSolver() : GpuSolver(SPHG_PCISPH) {}
};

class DFSPHSolver : public GpuSolver {   // src/DFSPHSolver.h:8-70
protected:
    Real maxErrorV = 0.1;
    Real cflFactor = 1.0;
    unsigned int minIterationsV = 1, maxIterationsV = 100, iterationsV = 0;

public:
    DFSPHSolver() : GpuSolver(SPHG_DFSPH) { minIterations = 2; maxIterations = 100; }
    void setCFLFactor(Real f) { cflFactor = f; }
    void setMaxErrorV(Real e) { maxErrorV = e; }
    void setMinIterationsV(unsigned int it) { minIterationsV = it; }
    void setMaxIterationsV(unsigned int it) { maxIterationsV = it; }
    Real getMaxErrorV() const { return maxErrorV; }
    unsigned int getMinIterationsV() const { return minIterationsV; }
    unsigned int getMaxIterationsV() const { return maxIterationsV; }
    void fillDesc(sphg_scene_desc& d) const override;
};

// ------------------------------------------------------- time manager

class TimeManager {   // src/TimeManager.h:15-117
    Real fps = 25, lastTimeSave = 0, time = 0, startTime = 0, endTime = 0, ts = 0.0001;
    Real minTimeStep = 0, maxTimeStep = 0;
    unsigned int frame = 1;
    std::map<std::string, std::chrono::steady_clock::time_point> starts;
    std::map<std::string, double> intervals;

public:
    void setFrame(unsigned int f) { frame = f; }
    void setTime(Real t) { time = t; }
    void setStartTime(Real t) { startTime = t; }
    void setEndTime(Real t) { endTime = t; }
    void setTimeStep(Real t) { ts = t; }
    void setFPS(Real f) { fps = f; }
    // min(1/fps, minTs) (src/TimeManager.h:55); call setFPS first
    void setMinTimeStep(Real m) { Real f = 1.0 / fps; minTimeStep = f < m ? f : m; }
    void setMaxTimeStep(Real m) { maxTimeStep = m; }
    unsigned int getFrame() const { return frame; }
    Real getTime() const { return time; }
    Real getStartTime() const { return startTime; }
    Real getEndTime() const { return endTime; }
    Real getTimeStep() const { return ts; }
    Real getFPS() const { return fps; }
    Real getMinTimeStep() const { return minTimeStep; }
    Real getMaxTimeStep() const { return maxTimeStep; }
    void startCounting(const std::string& n) { starts[n] = std::chrono::steady_clock::now(); }
    void stopCounting(const std::string& n) {
        intervals[n] = std::chrono::duration<double>(std::chrono::steady_clock::now() - starts[n]).count();
    }
    Real getInterval(const std::string& n) { return intervals[n]; }
    // the time at which hasToSave turns true
    Real nextSaveTime() const { return lastTimeSave + 1.0 / fps; }
    // src/TimeManager.h:101-111
    bool hasToSave() {
        if (time - lastTimeSave >= 1.0 / fps) {
            lastTimeSave = (Real)frame / fps;
            return true;
        }
        return false;
    }
};

// --------------------------------------------------------- scene data

struct SceneInfo {   // src/Simulation.h:47-68 (+ device extensions)
    Real startTime = 0, endTime = 1, timeStep = 1e-3, fps = 25, minTimeStep = 1e-6, maxTimeStep = 5e-3;
    Vector3r gravity{0, 0, -9.81};
    Real particleRadius = 0.025;
    int simulationMethod = 2, boundaryMethod = 2;
    Real stiffness = 1000, gamma = 7, eta = 0.01, etaV = 0.1, cflFactor = 1.0;
    unsigned int minIterations = 2, maxIterations = 100, minIterationsV = 1, maxIterationsV = 100;
    std::string device = "gpu";
    unsigned int maxNeighbors = 0;
};

struct BlockInfo {
    Vector3r min, max;
    Real jitter = 0;          // scene extension: +-jitter * 2r per coordinate
    unsigned long long seed = 0;
};

struct SphereInfo {           // scene extension: lattice sites inside a sphere
    Vector3r center;
    Real radius = 0;
};

struct EmitterInfo {   // src/Simulation.h:76-87
    unsigned int type = 0, numParticles = 0;
    Vector3r r;
    Real v = 0;
    Quat4r rot;
    Real startTime = 0, width = 0, height = 0, spacing = 0;
};

struct Fluid {   // src/Simulation.h:87-98
    std::vector<BlockInfo> fluidBlocks;
    std::vector<SphereInfo> fluidSpheres;
    Vector3r sphereLatticeMin, sphereLatticeMax;
    Real viscosity = 0, boundaryViscosity = 0, surfaceTension = 0, adhesion = 0, density0 = 1000;
    std::vector<EmitterInfo> emitters;
};

struct FluidInfo {
    std::vector<Fluid> fluids;
    int viscosityMethod = -1, surfaceTensionMethod = -1, adhesionMethod = -1;
};

struct Boundary {   // src/Simulation.h:119-127
    std::vector<std::pair<BlockInfo, bool>> box;
    std::vector<std::pair<SphereInfo, bool>> sphere;
    int numGeometry = 0;
    Real normalFct = 0, tangFct = 1;   // PCISPH boundary method
};

struct BoundaryInfo {
    std::vector<Boundary> boundaries;
};

struct SimulationInfo {
    SceneInfo sceneData;
    FluidInfo fluidData;
    BoundaryInfo boundaryData;
};

// ---------------------------------------------------------- simulation

class Simulation {   // src/Simulation.h:148-317
    static Simulation* current;
    std::string name;
    std::vector<std::unique_ptr<FluidModel>> fluidModels;
    std::vector<std::unique_ptr<BoundaryModel>> boundaryModels;
    std::unique_ptr<SPHSolver> solver;
    TimeManager tm;
    Real particleRadius = 0.025, supportRadius = 0.1;
    Vector3r gravity{0, 0, -9.81};
    int current_method = -1, current_bh_method = 0;
    int current_visco_method = -1, current_surften_method = -1, current_adhesion_method = -1;
    bool activeSave = false;
    unsigned int maxNeighbors = 0;
    int device = 0;

public:
    static const int WCSPH_METHOD, PCISPH_METHOD, DFSPH_METHOD;
    static const int CUBE_BOUNDARY_METHOD, PCISPH_BOUNDARY_METHOD, AKINCI_BOUNDARY_METHOD;
    static const int STANDARD_VISCOSITY_METHOD, ARTIFICIAL_VISCOSITY_METHOD, XSPH_VISCOSITY_METHOD;
    static const int AKINCI_SURFACE_TENSION_METHOD, AKINCI_ADHESION_METHOD;

    static Simulation* getCurrent();
    static void setCurrent(Simulation* s);
    static bool hasCurrent();
    static void destroyCurrent();

    void init();                 // src/Simulation.cpp:102-110
    bool step();                 // src/Simulation.cpp:148-182
    void run(unsigned int maxSteps = 0);   // steps while startTime <= time <= endTime

    void setParticleRadius(Real r) { particleRadius = r; supportRadius = 4.0 * r; }   // :192-199
    void setSupportRadius(Real h) { supportRadius = h; }
    void setGravity(Vector3r g) { gravity = g; }
    void setTime(Real t) { tm.setTime(t); }
    void setStartTime(Real t) { tm.setStartTime(t); }
    void setEndTime(Real t) { tm.setEndTime(t); }
    void setTimeStep(Real t) { tm.setTimeStep(t); }
    void setFPS(Real f) { tm.setFPS(f); }
    void setMinTimeStep(Real t) { tm.setMinTimeStep(t); }
    void setMaxTimeStep(Real t) { tm.setMaxTimeStep(t); }
    void setSimulationMethod(int method);   // src/Simulation.cpp:207-231
    void setBoundaryMethod(int m) { current_bh_method = m; }
    void setViscosityMethod(int m) { current_visco_method = m; }
    void setSurfaceTensionMethod(int m) { current_surften_method = m; }
    void setAdhesionMethod(int m) { current_adhesion_method = m; }
    void setMaxNeighbors(unsigned int n) { maxNeighbors = n; }
    void setDevice(int d) { device = d; }
    void activateSave(bool a) { activeSave = a; }

    Real getParticleRadius() const { return particleRadius; }
    Real getSupportRadius() const { return supportRadius; }
    Vector3r getGravity() const { return gravity; }
    Real getTime() const { return tm.getTime(); }
    unsigned int getFrame() const { return tm.getFrame(); }
    Real getStartTime() const { return tm.getStartTime(); }
    Real getEndTime() const { return tm.getEndTime(); }
    Real getTimeStep() const { return tm.getTimeStep(); }
    Real getFPS() const { return tm.getFPS(); }
    Real getMinTimeStep() const { return tm.getMinTimeStep(); }
    Real getMaxTimeStep() const { return tm.getMaxTimeStep(); }
    TimeManager& timeManager() { return tm; }
    int getSimulationMethod() const { return current_method; }
    int getBoundaryHandlingMethod() const { return current_bh_method; }
    int getViscosityMethod() const { return current_visco_method; }
    int getSurfaceTensionMethod() const { return current_surften_method; }
    int getAdhesionMethod() const { return current_adhesion_method; }
    unsigned int getMaxNeighbors() const { return maxNeighbors; }
    int getDevice() const { return device; }
    const std::string& getName() const { return name; }
    SPHSolver* getSolver() { return solver.get(); }

    FluidModel* addFluidModel(std::vector<Vector3r>& points, std::vector<Vector3r>& velocities);
    FluidModel* addFluidModel();
    FluidModel* getFluidModel(unsigned int i) { return fluidModels[i].get(); }
    unsigned int numberFluidModels() const { return (unsigned int)fluidModels.size(); }
    void addBoundaryModel(BoundaryModel* bm) { boundaryModels.emplace_back(bm); }
    BoundaryModel* getBoundaryModel(unsigned int i) { return boundaryModels[i].get(); }
    unsigned int numberBoundaryModels() const { return (unsigned int)boundaryModels.size(); }

    void initMasses();                     // src/Simulation.cpp:134-146
    void computeNonPressureForces();       // src/Simulation.cpp:313-323
    bool importScene(const std::string& path);   // src/Simulation.cpp:331-494
    // buildFluidBlock (src/Simulation.cpp:512-539) with the jitter extension.
    FluidModel* buildFluidBlock(const std::vector<BlockInfo>& blocks, const std::vector<SphereInfo>& spheres,
                                Vector3r sphereMin, Vector3r sphereMax);
    std::string describe();                // printInfo without the blocking getchar()
};

// --------------------------------------------------------- scene loader

class SceneLoader {   // src/SceneLoader.h
public:
    static unsigned int maxDigits;
    // src/SceneLoader.cpp:9-36; throws SphError with the offending key
    static bool readConfiguration(SimulationInfo& info, const std::string& path);
    static void readString(SimulationInfo& info, const std::string& text);
    // src/SceneLoader.cpp:204-268: <scene>_NNNN.json with numParticles,
    // numActiveParticles, time, timeStep, positions, velocities
    static std::string writeFluid();
    static std::string frameJson();
    // The frame text from the active particles' FP32 positions and
    // velocities; slots [active, total) are written as zeros.
    static std::string frameJsonFrom(unsigned int total, unsigned int active, Real time, Real timeStep,
                                     const float* x, const float* v);
};
