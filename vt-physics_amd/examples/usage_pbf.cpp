// usage_pbf - the programmatic set-up of the reference's Examples/PBF/Usage-1.cpp
// against this library's host API: a PBF solver from its template, a fluid
// cube and a boundary box built as objects, attached, then ticked.
//
//   usage_pbf [STEPS] [EXPORT_DIR]
#include <cstdio>
#include <cstdlib>

#include "../host/vt_physics.hpp"

using namespace VT_Physics;

int main(int argc, char **argv) {
    const unsigned steps = argc > 1 ? unsigned(std::atoi(argv[1])) : 10;
    const char *export_dir = argc > 2 ? argv[2] : nullptr;

    auto &sim = Simulator::getInstance();
    auto solverManager = sim.getSolverManager();
    auto objectManager = sim.getObjectManager();

    auto pbfSolver = solverManager->createSolver(PBF);
    auto pbf_config = pbfSolver->getSolverConfigTemplate();
    pbf_config["PBF"]["Required"]["animationTime"] = 5.0f;
    pbf_config["PBF"]["Required"]["timeStep"] = 0.004f;
    pbf_config["PBF"]["Required"]["particleRadius"] = 0.01f;
    pbf_config["PBF"]["Required"]["simSpaceLB"] = json({-0.2, -0.2, -0.2});
    pbf_config["PBF"]["Required"]["simSpaceSize"] = json({1.4, 1.4, 1.4});
    pbf_config["PBF"]["Required"]["iterationNum"] = 4;
    pbf_config["PBF"]["Required"]["XSPH_k"] = 0.01;
    pbf_config["PBF"]["Required"]["fPartRestDensity"] = 1000.f;
    pbf_config["PBF"]["Required"]["bPartRestDensity"] = 1000.f;
    if (export_dir) {
        pbf_config["EXPORT"]["Common"]["exportTargetDir"] = export_dir;
        pbf_config["EXPORT"]["SolverRequired"]["enable"] = true;
        pbf_config["EXPORT"]["SolverRequired"]["exportFps"] = 35;
    }
    if (!pbfSolver->setConfig(pbf_config)) return 1;

    auto pbfObjComponentConfig = pbfSolver->getSolverObjectComponentConfigTemplate();

    auto cube = objectManager->createObject(Particle_Cube);
    auto &cube_config = cube->getObjectComponentConfig();
    cube_config["epmMaterial"] = EPM_FLUID;
    cube_config["particleRadius"] = pbf_config["PBF"]["Required"]["particleRadius"];
    cube_config["lb"] = json({0.04, 0.04, 0.04});
    cube_config["size"] = json({0.4, 0.4, 0.4});
    auto cube_cfg = pbfObjComponentConfig;
    cube_cfg["exportFlag"] = true;
    cube_cfg["velocityStart"] = json({0.0, 0.0, 0.0});
    cube_cfg["colorStart"] = json({0.0, 1.0, 0.0});
    cube->attachSpecificSolverObjectComponentConfig(cube_cfg);
    cube->update();

    auto box = objectManager->createObject(Particle_Box);
    auto &box_config = box->getObjectComponentConfig();
    box_config["epmMaterial"] = EPM_BOUNDARY;
    box_config["particleRadius"] = pbf_config["PBF"]["Required"]["particleRadius"];
    box_config["lb"] = json({0.0, 0.0, 0.0});
    box_config["size"] = json({1.0, 1.0, 1.0});
    box_config["layerNum"] = 2;
    auto box_cfg = pbfObjComponentConfig;
    box_cfg["exportFlag"] = false;
    box->attachSpecificSolverObjectComponentConfig(box_cfg);
    box->update();

    if (!pbfSolver->attachObject(cube) || !pbfSolver->attachObject(box)) return 1;
    const bool ok = pbfSolver->tickNsteps(steps);
    std::printf("usage_pbf: %u ticks, %s\n", steps, ok ? "ok" : "failed");

    sim.terminate();
    Simulator::clean();
    return ok ? 0 : 1;
}
