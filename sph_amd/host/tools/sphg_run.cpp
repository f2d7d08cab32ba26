// sphg_run -- headless front end of the GPU solver: the reference's
// `-f scene.json [-o true|false]` command line (src/main.cpp:21-52,
// src/InputParser.h) without the GLUT window.
//
//   sphg_run -f scene.json [-o true|false] [--steps N] [--info] [--device D]
//            [--final state.json]
//
// --info parses and builds the scene and prints it without touching a GPU.
// Otherwise the scene is imported (Simulation::importScene), stepped until
// endTime (or N steps), writing <scene>_NNNN.json frames when -o true, and
// a final JSON line with steps, simulated time and device milliseconds.
// --final writes the last state as one frame file (the writeFluid format).
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "SPH.h"

namespace {

std::vector<std::string> tokens;

std::string option(const std::string& o) {
    for (size_t i = 0; i + 1 < tokens.size(); ++i)
        if (tokens[i] == o) return tokens[i + 1];
    return "";
}
bool has(const std::string& o) {
    for (auto& t : tokens)
        if (t == o) return true;
    return false;
}

}  // namespace

int main(int argc, char** argv) {
    for (int i = 1; i < argc; ++i) tokens.emplace_back(argv[i]);
    const std::string path = option("-f");
    if (path.empty()) {
        std::fprintf(stderr, "usage: sphg_run -f scene.json [-o true|false] [--steps N] [--info] [--device D] "
                             "[--final state.json]\n");
        return 2;
    }
    if (path.find("json") == std::string::npos) {
        std::fprintf(stderr, "File must be .json\n");
        return 2;
    }
    try {
        if (has("--info")) {
            SimulationInfo info;
            if (!SceneLoader::readConfiguration(info, path)) {
                std::fprintf(stderr, "Cannot read %s\n", path.c_str());
                return 1;
            }
            size_t blocks = 0, spheres = 0, boxes = 0, emitters = 0;
            unsigned long long budget = 0;
            for (auto& f : info.fluidData.fluids) {
                blocks += f.fluidBlocks.size();
                spheres += f.fluidSpheres.size();
                emitters += f.emitters.size();
                for (auto& e : f.emitters) budget += e.numParticles;
            }
            for (auto& b : info.boundaryData.boundaries) boxes += b.box.size();
            std::printf("{\"simulationMethod\": %d, \"boundaryMethod\": %d, \"particleRadius\": %.17g, "
                        "\"fluidModels\": %zu, \"fluidBlocks\": %zu, \"fluidSpheres\": %zu, \"boundaries\": %zu, "
                        "\"boxes\": %zu, \"viscosityMethod\": %d, \"surfaceTensionMethod\": %d, \"adhesionMethod\": %d, "
                        "\"emitters\": %zu, \"emitterParticles\": %llu}\n",
                        info.sceneData.simulationMethod, info.sceneData.boundaryMethod, info.sceneData.particleRadius,
                        info.fluidData.fluids.size(), blocks, spheres, info.boundaryData.boundaries.size(), boxes,
                        info.fluidData.viscosityMethod, info.fluidData.surfaceTensionMethod,
                        info.fluidData.adhesionMethod, emitters, budget);
            return 0;
        }
        Simulation* sim = Simulation::getCurrent();
        const std::string out = option("-o");
        sim->activateSave(out == "true");
        if (!option("--device").empty()) sim->setDevice(std::stoi(option("--device")));
        sim->importScene(path);
        std::fprintf(stderr, "%s", sim->describe().c_str());
        const unsigned int nsteps = option("--steps").empty() ? 0u : (unsigned int)std::stoul(option("--steps"));
        sim->run(nsteps);
        if (!option("--final").empty()) {
            std::ofstream f(option("--final"));
            if (!f.is_open()) throw SphError("cannot write " + option("--final"));
            f << SceneLoader::frameJson();
        }
        auto* gs = dynamic_cast<GpuSolver*>(sim->getSolver());
        const sphg_scene_scalars& s = gs->scalars();
        std::printf("{\"steps\": %u, \"time\": %.17g, \"time_step\": %.17g, \"frames\": %u, \"device_ms\": %.3f, "
                    "\"sum_iterations\": %u, \"sum_iterations_v\": %u}\n",
                    s.steps, s.time, s.time_step, sim->getFrame() - 1, gs->deviceMilliseconds(), s.sum_iterations,
                    s.sum_iterations_v);
        Simulation::destroyCurrent();
    } catch (const SphError& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
