// scene_runner - runs a scene file in the reference's format
// (Examples/<Solver>/Scene-1/Configuration.json: one solver domain, EXPORT
// and OBJECTS) through the host API, the way the reference's Usage-2
// examples do (Examples/PBF/Usage-2.cpp).
//
//   scene_runner SCENE.json [--steps N | --run] [--export-dir DIR]
//                [--dump FILE] [--generate-only DIR]
//
// --steps N          tickNsteps(N) (default: run() until animationTime)
// --export-dir DIR   overrides EXPORT.Common.exportTargetDir
// --dump FILE        final positions, float32 [n][3] in attach order
// --generate-only DIR  builds the objects only (no GPU) and writes each
//                    object's particles to DIR/<index>.bin as float32 [n][3]
// Prints one JSON line with the outcome.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../host/vt_physics.hpp"

using namespace VT_Physics;

static int usage() {
    std::fprintf(stderr,
                 "usage: scene_runner SCENE.json [--steps N | --run] [--export-dir DIR] [--dump FILE] "
                 "[--generate-only DIR]\n");
    return 2;
}

int main(int argc, char **argv) {
    if (argc < 2) return usage();
    std::string scene_path = argv[1], export_dir, dump, gen_dir;
    long steps = -1;
    for (int i = 2; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--steps" && i + 1 < argc)
            steps = std::atol(argv[++i]);
        else if (a == "--run")
            steps = -1;
        else if (a == "--export-dir" && i + 1 < argc)
            export_dir = argv[++i];
        else if (a == "--dump" && i + 1 < argc)
            dump = argv[++i];
        else if (a == "--generate-only" && i + 1 < argc)
            gen_dir = argv[++i];
        else
            return usage();
    }
    json scene = JsonHandler::loadJson(scene_path);
    if (scene.empty()) return 1;

    auto &sim = Simulator::getInstance();
    ObjectManager *objects = sim.getObjectManager();
    if (!gen_dir.empty()) {
        auto objs = objects->createObjectsFromJson(scene);
        std::printf("{\"objects\": [");
        for (size_t i = 0; i < objs.size(); ++i) {
            auto el = objs[i]->getObjectComponent() ? objs[i]->getObjectComponent()->getElements()
                                                    : std::vector<float>{};
            const std::string path = gen_dir + "/" + std::to_string(i) + ".bin";
            FILE *f = std::fopen(path.c_str(), "wb");
            if (!f) return 1;
            std::fwrite(el.data(), sizeof(float), el.size(), f);
            std::fclose(f);
            std::printf("%s%zu", i ? ", " : "", el.size() / 3);
        }
        std::printf("]}\n");
        sim.terminate();
        return objs.empty() ? 1 : 0;
    }

    int solver = -1;
    const char *domains[4] = {"PBF", "DFSPH", "IMM", "IMMCT"};
    for (int s = 0; s < 4; ++s)
        if (scene.contains(domains[s])) solver = s;
    if (solver < 0) {
        std::fprintf(stderr, "scene_runner: no solver domain (PBF/DFSPH/IMM/IMMCT) in %s\n", scene_path.c_str());
        return 1;
    }
    if (!export_dir.empty()) scene["EXPORT"]["Common"]["exportTargetDir"] = export_dir;

    Solver *solverPtr = sim.getSolverManager()->createSolver(SolverType(solver));
    bool ok = solverPtr->setConfig(scene);
    auto objs = objects->createObjectsFromJson(scene);
    ok = ok && solverPtr->attachObjects(objs);
    ok = ok && (steps >= 0 ? solverPtr->tickNsteps(uint32_t(steps)) : solverPtr->run());

    auto *ps = static_cast<ParticleSolver *>(solverPtr);
    vtp_step_stats st{};
    if (ps->context()) vtp_stats(ps->context(), &st);
    if (ok && !dump.empty()) {
        std::vector<float> pos;
        if (!ps->readField(VTP_FIELD_POS, pos)) {
            ok = false;
        } else {
            FILE *f = std::fopen(dump.c_str(), "wb");
            if (!f) return 1;
            std::fwrite(pos.data(), sizeof(float), pos.size(), f);
            std::fclose(f);
        }
    }
    std::printf("{\"solver\": \"%s\", \"ok\": %s, \"particles\": %u, \"steps\": %llu, \"sim_time\": %.9g, "
                "\"crash_flags\": %u, \"dead_particles\": %u}\n",
                domains[solver], ok ? "true" : "false", ps->particleNum(), (unsigned long long)st.steps,
                double(st.sim_time), st.crash_flags, st.dead_particles);
    sim.terminate();
    Simulator::clean();
    return ok ? 0 : 1;
}
