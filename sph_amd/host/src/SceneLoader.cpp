// SceneLoader.cpp -- the reference's scene JSON format
// (src/SceneLoader.cpp:38-202) and frame output (:204-268), plus
// Simulation::importScene (src/Simulation.cpp:331-494).
//
// Keys read exactly as the reference does: every key it reads is required
// (a missing key is an error naming it).  Additive keys, ignored by the
// reference (it reads only the keys above):
//   Configuration.device          "gpu" (default)
//   Configuration.maxNeighbors    neighbor-list capacity (0 = library default)
//   FluidModel[i].fluidBlocks[j].jitter / .seed   seeded lattice jitter
//   FluidModel[i].fluidSpheres[j] {center[3], radius}, with
//   FluidModel[i].sphereLattice {min[3], max[3]}  lattice of the spheres
#include <fstream>
#include <sstream>

#include "Json.h"
#include "SPH.h"

using sph::Json;
using sph::JsonError;

unsigned int SceneLoader::maxDigits = 4;

namespace {

Vector3r vec3(const Json& j, const char* what) {
    if (!j.isArray() || j.size() < 3) throw JsonError(std::string(what) + ": expected [x, y, z]");
    return Vector3r(j.at((size_t)0).num(what), j.at((size_t)1).num(what), j.at((size_t)2).num(what));
}

// src/SceneLoader.cpp:38-76
void readSceneInfo(SimulationInfo& info, const Json& cfgRoot) {
    const Json& c = cfgRoot.at("Configuration");
    SceneInfo& s = info.sceneData;
    s.startTime = c.at("startTime").num("startTime");
    s.endTime = c.at("endTime").num("endTime");
    s.timeStep = c.at("timeStep").num("timeStep");
    s.fps = c.at("fps").num("fps");
    s.minTimeStep = c.at("minTimeStep").num("minTimeStep");
    s.maxTimeStep = c.at("maxTimeStep").num("maxTimeStep");
    s.cflFactor = c.at("cflFactor").num("cflFactor");
    s.particleRadius = c.at("particleRadius").num("particleRadius");
    s.boundaryMethod = c.at("boundaryMethod").integer("boundaryMethod");
    s.simulationMethod = c.at("simulationMethod").integer("simulationMethod");
    s.gravity = vec3(c.at("gravity"), "gravity");
    if (s.simulationMethod == Simulation::WCSPH_METHOD) {
        s.stiffness = c.at("stiffness").num("stiffness");
        s.gamma = c.at("gamma").num("gamma");
    } else if (s.simulationMethod == Simulation::PCISPH_METHOD) {
        s.eta = c.at("eta").num("eta");
        s.minIterations = (unsigned)c.at("minIterations").integer("minIterations");
        s.maxIterations = (unsigned)c.at("maxIterations").integer("maxIterations");
    } else if (s.simulationMethod == Simulation::DFSPH_METHOD) {
        s.eta = c.at("eta").num("eta");
        s.etaV = c.at("etaV").num("etaV");
        s.minIterations = (unsigned)c.at("minIterations").integer("minIterations");
        s.maxIterations = (unsigned)c.at("maxIterations").integer("maxIterations");
        s.minIterationsV = (unsigned)c.at("minIterationsV").integer("minIterationsV");
        s.maxIterationsV = (unsigned)c.at("maxIterationsV").integer("maxIterationsV");
    }
    if (c.has("device")) s.device = c.at("device").str("device");
    if (c.has("maxNeighbors")) s.maxNeighbors = (unsigned)c.at("maxNeighbors").integer("maxNeighbors");
}

// src/SceneLoader.cpp:78-141
void readFluidInfo(SimulationInfo& info, const Json& root) {
    const Json& f = root.at("Fluid");
    FluidInfo& fi = info.fluidData;
    fi.viscosityMethod = f.at("viscosityMethod").integer("viscosityMethod");
    fi.surfaceTensionMethod = f.at("surfaceTensionMethod").integer("surfaceTensionMethod");
    fi.adhesionMethod = f.at("adhesionMethod").integer("adhesionMethod");
    const Json& models = f.get("FluidModel");
    fi.fluids.resize(models.size());
    for (size_t i = 0; i < models.size(); ++i) {
        const Json& m = models.at(i);
        Fluid& fl = fi.fluids[i];
        fl.density0 = m.at("density0").num("density0");
        if (fi.viscosityMethod > -1) {
            fl.viscosity = m.at("viscosity").num("viscosity");
            fl.boundaryViscosity = m.at("boundaryViscosity").num("boundaryViscosity");
        }
        if (fi.surfaceTensionMethod > -1) fl.surfaceTension = m.at("surfaceTension").num("surfaceTension");
        if (fi.adhesionMethod > -1) fl.adhesion = m.at("adhesion").num("adhesion");
        const Json& blocks = m.get("fluidBlocks");
        for (size_t j = 0; j < blocks.size(); ++j) {
            const Json& b = blocks.at(j);
            BlockInfo bi;
            bi.min = vec3(b.at("min"), "fluidBlocks.min");
            bi.max = vec3(b.at("max"), "fluidBlocks.max");
            if (b.has("jitter")) bi.jitter = b.at("jitter").num("jitter");
            if (b.has("seed")) bi.seed = (unsigned long long)b.at("seed").num("seed");
            fl.fluidBlocks.push_back(bi);
        }
        const Json& spheres = m.get("fluidSpheres");
        for (size_t j = 0; j < spheres.size(); ++j) {
            SphereInfo s;
            s.center = vec3(spheres.at(j).at("center"), "fluidSpheres.center");
            s.radius = spheres.at(j).at("radius").num("fluidSpheres.radius");
            fl.fluidSpheres.push_back(s);
        }
        if (spheres.size()) {
            const Json& lat = m.at("sphereLattice");
            fl.sphereLatticeMin = vec3(lat.at("min"), "sphereLattice.min");
            fl.sphereLatticeMax = vec3(lat.at("max"), "sphereLattice.max");
        }
        // src/SceneLoader.cpp:107-139
        const Json& ems = m.get("emitters");
        for (size_t j = 0; j < ems.size(); ++j) {
            const Json& e = ems.at(j);
            EmitterInfo ei;
            ei.r = vec3(e.at("position"), "emitters.position");
            const Json& rot = e.at("rotation");
            if (rot.size() != 4) throw SphError("emitters.rotation: expected 4 numbers (w, x, y, z)");
            ei.rot = Quat4r(rot.at(0).num("emitters.rotation"), rot.at(1).num("emitters.rotation"),
                            rot.at(2).num("emitters.rotation"), rot.at(3).num("emitters.rotation"));
            ei.v = e.at("velocity").num("emitters.velocity");
            ei.numParticles = (unsigned int)e.at("numParticles").num("emitters.numParticles");
            ei.startTime = e.at("startTime").num("emitters.startTime");
            ei.type = (unsigned int)e.at("type").num("emitters.type");
            ei.width = e.at("width").num("emitters.width");
            ei.height = e.at("height").num("emitters.height");
            ei.spacing = e.at("spacing").num("emitters.spacing");
            fl.emitters.push_back(ei);
        }
    }
}

// src/SceneLoader.cpp:143-202
void readBoundaryInfo(SimulationInfo& info, const Json& root) {
    const Json& arr = root.get("Boundary");
    BoundaryInfo& bi = info.boundaryData;
    bi.boundaries.resize(arr.size());
    const bool pcisphBoundary = info.sceneData.boundaryMethod == Simulation::PCISPH_BOUNDARY_METHOD;
    for (size_t i = 0; i < arr.size(); ++i) {
        const Json& b = arr.at(i);
        Boundary& bd = bi.boundaries[i];
        const Json& boxes = b.get("box");
        for (size_t j = 0; j < boxes.size(); ++j) {
            BlockInfo box;
            box.min = vec3(boxes.at(j).at("min"), "box.min");
            box.max = vec3(boxes.at(j).at("max"), "box.max");
            bool inverted = pcisphBoundary ? boxes.at(j).at("inverted").flag("inverted") : false;
            bd.box.emplace_back(box, inverted);
        }
        const Json& spheres = b.get("sphere");
        for (size_t j = 0; j < spheres.size(); ++j) {
            SphereInfo s;
            s.center = vec3(spheres.at(j).at("pos"), "sphere.pos");
            s.radius = spheres.at(j).at("radius").num("sphere.radius");
            bool inverted = pcisphBoundary ? spheres.at(j).at("inverted").flag("inverted") : false;
            bd.sphere.emplace_back(s, inverted);
        }
        if (pcisphBoundary) {   // src/SceneLoader.cpp:161-165
            bd.normalFct = b.at("normalFct").num("Boundary.normalFct");
            bd.tangFct = b.at("tangFct").num("Boundary.tangFct");
        }
        bd.numGeometry = (int)b.get("geometry").size();
    }
}

void readAll(SimulationInfo& info, const Json& root) {
    try {
        readSceneInfo(info, root);
        readFluidInfo(info, root);
        readBoundaryInfo(info, root);
    } catch (const JsonError& e) {
        throw SphError(std::string("scene: ") + e.what());
    }
}

}  // namespace

void SceneLoader::readString(SimulationInfo& info, const std::string& text) {
    Json root;
    try {
        root = Json::parse(text);
    } catch (const JsonError& e) {
        throw SphError(std::string("scene: ") + e.what());
    }
    readAll(info, root);
}

bool SceneLoader::readConfiguration(SimulationInfo& info, const std::string& path) {
    std::ifstream f(path);
    if (!f.is_open()) return false;   // the reference logs "Cannot read" and returns false
    std::stringstream ss;
    ss << f.rdbuf();
    readString(info, ss.str());
    return true;
}

std::string SceneLoader::frameJson() {
    Simulation* sim = Simulation::getCurrent();
    if (auto* gs = dynamic_cast<GpuSolver*>(sim->getSolver())) gs->syncToHost();
    Json state = Json::object();
    unsigned int total = 0, active = 0;
    for (unsigned int i = 0; i < sim->numberFluidModels(); ++i) {
        total += sim->getFluidModel(i)->getNumParticles();
        active += sim->getFluidModel(i)->getNumActiveParticles();
    }
    state["numParticles"] = Json::number(total);
    state["numActiveParticles"] = Json::number(active);
    state["time"] = Json::number(sim->getTime());
    state["timeStep"] = Json::number(sim->getTimeStep());
    Json pos = Json::array(), vel = Json::array();
    for (unsigned int f = 0; f < sim->numberFluidModels(); ++f) {
        FluidModel* fm = sim->getFluidModel(f);
        for (unsigned int i = 0; i < fm->getNumParticles(); ++i) {
            const Vector3r& p = fm->getPosition(i);
            const Vector3r& v = fm->getVelocity(i);
            Json a = Json::array(), b = Json::array();
            a.push(Json::number(p.x)); a.push(Json::number(p.y)); a.push(Json::number(p.z));
            b.push(Json::number(v.x)); b.push(Json::number(v.y)); b.push(Json::number(v.z));
            pos.push(a);
            vel.push(b);
        }
    }
    state["positions"] = pos;
    state["velocities"] = vel;
    return state.dump();
}

std::string SceneLoader::frameJsonFrom(unsigned int total, unsigned int active, Real time, Real timeStep,
                                       const float* x, const float* v) {
    Json state = Json::object();
    state["numParticles"] = Json::number(total);
    state["numActiveParticles"] = Json::number(active);
    state["time"] = Json::number(time);
    state["timeStep"] = Json::number(timeStep);
    Json pos = Json::array(), vel = Json::array();
    for (unsigned int i = 0; i < total; ++i) {
        Json a = Json::array(), b = Json::array();
        for (int c = 0; c < 3; ++c) {
            a.push(Json::number(i < active ? (double)x[3 * i + c] : 0.0));
            b.push(Json::number(i < active ? (double)v[3 * i + c] : 0.0));
        }
        pos.push(a);
        vel.push(b);
    }
    state["positions"] = pos;
    state["velocities"] = vel;
    return state.dump();
}

// <scene>_NNNN.json next to the scene file (src/SceneLoader.cpp:204-268).
std::string SceneLoader::writeFluid() {
    Simulation* sim = Simulation::getCurrent();
    std::string filename = sim->getName();
    std::string num = std::to_string(sim->getFrame());
    while (num.size() < maxDigits) num = "0" + num;
    const std::string ext = ".json";
    if (filename.size() >= ext.size() && filename.compare(filename.size() - ext.size(), ext.size(), ext) == 0)
        filename.insert(filename.size() - ext.size(), "_" + num);
    else
        filename += "_" + num + ext;
    if (auto* gs = dynamic_cast<GpuSolver*>(sim->getSolver())) {
        gs->queueFrame(filename);   // written beside the next steps
        return filename;
    }
    std::ofstream f(filename);
    if (!f.is_open()) throw SphError("cannot write " + filename);
    f << frameJson();
    return filename;
}

// src/Simulation.cpp:331-494
bool Simulation::importScene(const std::string& path) {
    name = path;
    SimulationInfo info;
    if (!SceneLoader::readConfiguration(info, path)) throw SphError("cannot read " + path);
    const SceneInfo& s = info.sceneData;
    if (s.device != "gpu") throw SphError("Configuration.device '" + s.device + "': sph_amd runs on the GPU only");
    setTimeStep(s.timeStep);
    setFPS(s.fps);
    setMinTimeStep(s.minTimeStep);
    setMaxTimeStep(s.maxTimeStep);
    setStartTime(s.startTime);
    setEndTime(s.endTime);
    setSimulationMethod(s.simulationMethod);
    setBoundaryMethod(s.boundaryMethod);
    setGravity(s.gravity);
    setParticleRadius(s.particleRadius);
    setMaxNeighbors(s.maxNeighbors);
    if (s.simulationMethod == WCSPH_METHOD) {
        auto* w = static_cast<WCSPHSolver*>(getSolver());
        w->setStiffness(s.stiffness);
        w->setGamma(s.gamma);
    } else if (s.simulationMethod == PCISPH_METHOD) {
        auto* p = static_cast<PCISPHSolver*>(getSolver());
        p->setMaxError(s.eta);
        p->setMinIterations(s.minIterations);
        p->setMaxIterations(s.maxIterations);
    } else {
        auto* d = static_cast<DFSPHSolver*>(getSolver());
        d->setMaxError(s.eta);
        d->setMaxErrorV(s.etaV);
        d->setCFLFactor(s.cflFactor);
        d->setMinIterations(s.minIterations);
        d->setMaxIterations(s.maxIterations);
        d->setMinIterationsV(s.minIterationsV);
        d->setMaxIterationsV(s.maxIterationsV);
    }
    const FluidInfo& fi = info.fluidData;
    setViscosityMethod(fi.viscosityMethod);
    setSurfaceTensionMethod(fi.surfaceTensionMethod);
    setAdhesionMethod(fi.adhesionMethod);
    for (const Fluid& fl : fi.fluids) {
        FluidModel* fm = nullptr;
        if (!fl.fluidBlocks.empty() || !fl.fluidSpheres.empty())
            fm = buildFluidBlock(fl.fluidBlocks, fl.fluidSpheres, fl.sphereLatticeMin, fl.sphereLatticeMax);
        if (!fm) fm = addFluidModel();
        for (const EmitterInfo& e : fl.emitters)   // src/Simulation.cpp:405-417
            fm->addEmitter(e.type, e.numParticles, e.r, e.v, e.rot, e.startTime, e.width, e.height, e.spacing);
        fm->setRefDensity(fl.density0);
        if (fi.viscosityMethod > -1 && fl.viscosity != 0.0) fm->setViscosityForce(fl.viscosity, fl.boundaryViscosity);
        if (fi.surfaceTensionMethod > -1 && fl.surfaceTension != 0.0) fm->setSurfaceTensionForce(fl.surfaceTension);
        if (fi.adhesionMethod > -1 && fl.adhesion != 0.0) fm->setAdhesionForce(fl.adhesion);
    }
    if (s.boundaryMethod == AKINCI_BOUNDARY_METHOD) {
        for (const Boundary& b : info.boundaryData.boundaries) {
            if (b.numGeometry) throw SphError("OBJ geometry boundaries are not supported by the GPU solver");
            auto* bm = new AkinciBoundaryModel();
            addBoundaryModel(bm);   // owned before sampling, so a throw cannot leak it
            for (auto& box : b.box) bm->addCube(box.first.min, box.first.max);
            for (auto& sp : b.sphere) bm->addSphere(sp.first.center, sp.first.radius);
        }
    } else if (s.boundaryMethod == PCISPH_BOUNDARY_METHOD) {
        // src/Simulation.cpp:459-486: one model per Boundary entry, boundary
        // mass rho0 (2r)^3 and rest density of the (single) fluid
        if (numberFluidModels() != 1) throw SphError("the PCISPH boundary method takes one fluid model");
        const Real rho0 = getFluidModel(0)->getRefDensity();
        for (const Boundary& b : info.boundaryData.boundaries) {
            if (b.numGeometry) throw SphError("OBJ geometry boundaries are not supported by the GPU solver");
            auto* bm = new PCISPHBoundaryModel();
            addBoundaryModel(bm);
            bm->setMass(rho0 * std::pow(2.0 * getParticleRadius(), 3.0));
            bm->setRefDensity(rho0);
            bm->setNormalFct(b.normalFct);
            bm->setTangentialFct(b.tangFct);
            for (auto& box : b.box) bm->addCube(box.first.min, box.first.max, box.second);
            for (auto& sp : b.sphere) bm->addSphere(sp.first.center, sp.first.radius, sp.second);
        }
    } else if (s.boundaryMethod != CUBE_BOUNDARY_METHOD) {
        // cube (0): the reference's loader creates no boundary model
        // (src/Simulation.cpp:435-486)
        throw SphError("boundaryMethod " + std::to_string(s.boundaryMethod) +
                       " is not one of 0 (cube), 1 (PCISPH), 2 (Akinci)");
    }
    init();
    return true;
}
