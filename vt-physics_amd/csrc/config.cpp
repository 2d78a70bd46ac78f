// config.cpp - solver JSON configuration -> vtp_params.
//
// Key names, the Required/Optional split and the template defaults are the
// reference's config API (ConfigTemplates/<Solver>/<Solver>ConfigTemplate.json,
// required-key lists in PBFSolver.hpp:18-34, DFSPHSolver.hpp:18-36,
// IMMSolver.hpp:17-37, IMMCTSolver.hpp:17-43); they are kept identical so a
// reference scene file configures this library unchanged.  Derived values
// follow PBFSolver.cpp:43-58 (gravity (0,-9.8,0) unless Optional.enable).
//
// Extension keys (no reference counterpart) live in an optional
// "<Solver>.Extensions" object: cflEnable, cflNumber, dtMax,
// maxSolverIterations, polymerPhase, phaseRestViscosity (IMM-CT only),
// massWeighted (neighbour mass in the PBF gradient / IMM-CT stress term).
#include <cstring>
#include <string>

#include "../../include/vtp.h"
#include "../host/json.hpp"

namespace {

using vtp::Json;

const char *kDomain[4] = {"PBF", "DFSPH", "IMM", "IMMCT"};

const std::vector<std::string> kRequired[4] = {
    {"animationTime", "timeStep", "particleRadius", "simSpaceLB", "simSpaceSize", "maxNeighborNum",
     "iterationNum", "XSPH_k", "fPartRestDensity", "bPartRestDensity"},
    {"animationTime", "timeStep", "particleRadius", "simSpaceLB", "simSpaceSize", "maxNeighborNum",
     "fPartRestDensity", "bPartRestDensity", "fPartRestViscosity", "divFreeThreshold", "incompThreshold",
     "surfaceTensionCoefficient"},
    {"animationTime", "timeStep", "particleRadius", "simSpaceLB", "simSpaceSize", "maxNeighborNum",
     "divFreeThreshold", "incompThreshold", "surfaceTensionCoefficient", "diffusionCoefficientCf",
     "momentumExchangeCoefficientCd", "phaseRestDensity", "phaseRestViscosity", "phaseRestColor"},
    {"animationTime", "timeStep", "particleRadius", "simSpaceLB", "simSpaceSize", "maxNeighborNum",
     "divFreeThreshold", "incompThreshold", "surfaceTensionCoefficient", "diffusionCoefficientCf",
     "momentumExchangeCoefficientCd0", "solventViscosity", "phaseRestDensity", "phaseRestColor",
     "solutionBasicViscosity", "solutionMaxViscosity", "relaxationTime", "shearThinningBasicFactor",
     "rheologicalThreshold"},
};
const std::vector<std::string> kOptional = {"enable", "gravity"};

// Template defaults (same values as the reference's template files).
const char *kTemplate[4] = {
    R"({"PBF":{"Required":{"animationTime":10,"timeStep":0.01,"particleRadius":0.075,
        "simSpaceLB":[-1,-1,-1],"simSpaceSize":[2,2,2],"maxNeighborNum":60,"iterationNum":10,
        "XSPH_k":0.02,"fPartRestDensity":1000,"bPartRestDensity":1000},
        "Optional":{"enable":false,"gravity":[0,-9.81,0]}},
       "EXPORT":{"Common":{"exportTargetDir":"path/to/output","exportFilePrefix":"obj","exportFileType":"PLY"},
        "SolverRequired":{"enable":false,"exportFps":35,"exportGroupPolicy":"MERGE"}}})",
    R"({"DFSPH":{"Required":{"animationTime":10,"timeStep":0.01,"particleRadius":0.075,
        "simSpaceLB":[-1,-1,-1],"simSpaceSize":[2,2,2],"maxNeighborNum":60,"fPartRestDensity":1000,
        "bPartRestDensity":1000,"fPartRestViscosity":0.001,"divFreeThreshold":1e-4,"incompThreshold":1e-4,
        "surfaceTensionCoefficient":0.001},
        "Optional":{"enable":false,"gravity":[0,-9.81,0]}},
       "EXPORT":{"Common":{"exportTargetDir":"path/to/output","exportFilePrefix":"obj","exportFileType":"PLY"},
        "SolverRequired":{"enable":false,"exportFps":35,"exportGroupPolicy":"MERGE"}}})",
    R"({"IMM":{"Required":{"animationTime":10,"timeStep":0.01,"particleRadius":0.075,
        "simSpaceLB":[-10,-10,-10],"simSpaceSize":[20,20,20],"maxNeighborNum":60,"divFreeThreshold":1e-4,
        "incompThreshold":1e-4,"surfaceTensionCoefficient":0.001,"diffusionCoefficientCf":0.1,
        "momentumExchangeCoefficientCd":0.5,"phaseRestDensity":[1000,1000,1000],
        "phaseRestViscosity":[0.01,0.01,0.01],"phaseRestColor":[[255,0,0],[0,255,0],[0,0,255]]},
        "Optional":{"enable":false,"gravity":[0,-9.81,0]}},
       "EXPORT":{"Common":{"exportTargetDir":"path/to/output","exportFilePrefix":"obj","exportFileType":"PLY"},
        "SolverRequired":{"enable":false,"exportFps":35,"exportGroupPolicy":"MERGE"}}})",
    R"({"IMMCT":{"Required":{"animationTime":5,"timeStep":0.0008,"particleRadius":0.05,
        "simSpaceLB":[-10,-10,-10],"simSpaceSize":[20,20,20],"maxNeighborNum":60,"divFreeThreshold":1e-4,
        "incompThreshold":1e-4,"surfaceTensionCoefficient":0.001,"diffusionCoefficientCf":0.15,
        "momentumExchangeCoefficientCd0":0.5,"solventViscosity":0.01,"phaseRestDensity":[900,1000],
        "phaseRestColor":[[255,0,0],[0,0,255]],"solutionBasicViscosity":12,"solutionMaxViscosity":12,
        "relaxationTime":0.001,"shearThinningBasicFactor":0.5,"rheologicalThreshold":0.4},
        "Optional":{"enable":false,"gravity":[0,-9.81,0]}},
       "EXPORT":{"Common":{"exportTargetDir":"path/to/output","exportFilePrefix":"obj","exportFileType":"PLY"},
        "SolverRequired":{"enable":false,"exportFps":35,"exportGroupPolicy":"MERGE"}}})",
};

void set_err(char *err, size_t len, const std::string &msg) {
    if (!err || len == 0) return;
    std::strncpy(err, msg.c_str(), len - 1);
    err[len - 1] = '\0';
}

void vec3(const Json &j, float out[3]) {
    auto v = j.floats();
    if (v.size() != 3) throw std::runtime_error("expected a 3-vector");
    out[0] = v[0];
    out[1] = v[1];
    out[2] = v[2];
}

}  // namespace

namespace vtp {
const char *solver_template(int solver) { return (solver >= 0 && solver < 4) ? kTemplate[solver] : nullptr; }
const char *solver_domain(int solver) { return (solver >= 0 && solver < 4) ? kDomain[solver] : nullptr; }
}  // namespace vtp

extern "C" vtp_status vtp_params_from_json(int32_t solver, const char *json_text, vtp_params *out, char *err,
                                           size_t err_len) {
    if (solver < 0 || solver > 3 || !json_text || !out) {
        set_err(err, err_len, "invalid argument");
        return VTP_EINVAL;
    }
    try {
        Json doc = Json::parse(json_text);
        const std::string dom = kDomain[solver];
        if (doc.empty()) throw std::runtime_error(dom + " config json is empty.");
        if (!doc.contains(dom)) throw std::runtime_error(dom + " config missing main domain: " + dom);
        const Json &d = doc.at(dom);
        if (!d.contains("Required")) throw std::runtime_error(dom + " config missing necessary domain: Required");
        const Json &r = d.at("Required");
        for (const auto &k : kRequired[solver])
            if (!r.contains(k)) throw std::runtime_error(dom + " config missing Required key: " + k);
        for (const auto &k : kOptional)
            if (!d.contains("Optional") || !d.at("Optional").contains(k))
                throw std::runtime_error(dom + " config missing Optional key: " + k);
        const Json &o = d.at("Optional");

        vtp_params p;
        std::memset(&p, 0, sizeof p);
        p.solver = solver;
        p.animation_time = float(r["animationTime"].number());
        p.time_step = float(r["timeStep"].number());
        p.particle_radius = float(r["particleRadius"].number());
        vec3(r["simSpaceLB"], p.sim_space_lb);
        vec3(r["simSpaceSize"], p.sim_space_size);
        p.max_neighbor_num = uint32_t(r["maxNeighborNum"].number());
        p.gravity[0] = 0.f;
        p.gravity[1] = -9.8f;
        p.gravity[2] = 0.f;
        if (o["enable"].boolean()) vec3(o["gravity"], p.gravity);
        p.max_solver_iterations = 100;
        p.polymer_phase = 1;
        p.cfl_number = 0.4f;
        p.dt_max = p.time_step;

        if (solver == VTP_PBF) {
            p.iteration_num = uint32_t(r["iterationNum"].number());
            p.xsph_k = float(r["XSPH_k"].number());
            p.f_rest_density = float(r["fPartRestDensity"].number());
            p.b_rest_density = float(r["bPartRestDensity"].number());
        } else if (solver == VTP_DFSPH) {
            p.f_rest_density = float(r["fPartRestDensity"].number());
            p.b_rest_density = float(r["bPartRestDensity"].number());
            p.f_rest_viscosity = float(r["fPartRestViscosity"].number());
        }
        if (solver != VTP_PBF) {
            p.div_free_threshold = float(r["divFreeThreshold"].number());
            p.incomp_threshold = float(r["incompThreshold"].number());
            p.surface_tension_coefficient = float(r["surfaceTensionCoefficient"].number());
        }
        if (solver == VTP_IMM || solver == VTP_IMMCT) {
            auto rho = r["phaseRestDensity"].floats();
            auto col = r["phaseRestColor"].float_rows();
            if (rho.empty() || rho.size() > VTP_MAX_PHASES)
                throw std::runtime_error(dom + " phaseRestDensity must hold 1.." +
                                         std::to_string(VTP_MAX_PHASES) + " phases");
            if (col.size() != rho.size())
                throw std::runtime_error(dom + " config phase-related parameter size not match");
            p.phase_num = uint32_t(rho.size());
            p.cf = float(r["diffusionCoefficientCf"].number());
            for (uint32_t k = 0; k < p.phase_num; ++k) {
                p.phase_rest_density[k] = rho[k];
                if (col[k].size() != 3) throw std::runtime_error(dom + " phaseRestColor rows must be 3-vectors");
                for (int a = 0; a < 3; ++a) p.phase_rest_color[k][a] = col[k][a];
            }
            if (solver == VTP_IMM) {
                auto vis = r["phaseRestViscosity"].floats();
                if (vis.size() != rho.size())
                    throw std::runtime_error(dom + " config phase-related parameter size not match");
                for (uint32_t k = 0; k < p.phase_num; ++k) p.phase_rest_viscosity[k] = vis[k];
                p.cd = float(r["momentumExchangeCoefficientCd"].number());
            } else {
                float sv = float(r["solventViscosity"].number());
                for (uint32_t k = 0; k < p.phase_num; ++k) p.phase_rest_viscosity[k] = sv;
                p.cd = float(r["momentumExchangeCoefficientCd0"].number());
                p.solution_vis_base = float(r["solutionBasicViscosity"].number());
                p.solution_vis_max = float(r["solutionMaxViscosity"].number());
                p.relaxation_time = float(r["relaxationTime"].number());
                p.thinning_exp0 = float(r["shearThinningBasicFactor"].number());
                p.rheology_threshold = float(r["rheologicalThreshold"].number());
            }
        }
        if (d.contains("Extensions")) {
            const Json &e = d.at("Extensions");
            if (e.contains("cflEnable")) p.cfl_enable = e["cflEnable"].boolean() ? 1 : 0;
            if (e.contains("cflNumber")) p.cfl_number = float(e["cflNumber"].number());
            if (e.contains("dtMax")) p.dt_max = float(e["dtMax"].number());
            if (e.contains("maxSolverIterations"))
                p.max_solver_iterations = uint32_t(e["maxSolverIterations"].number());
            if (e.contains("polymerPhase")) p.polymer_phase = int32_t(e["polymerPhase"].number());
            if (e.contains("massWeighted")) p.mass_weighted = e["massWeighted"].boolean() ? 1 : 0;
            if (e.contains("phaseRestViscosity") && solver == VTP_IMMCT) {
                auto vis = e["phaseRestViscosity"].floats();
                if (vis.size() != p.phase_num)
                    throw std::runtime_error(dom + " Extensions.phaseRestViscosity size not match");
                for (uint32_t k = 0; k < p.phase_num; ++k) p.phase_rest_viscosity[k] = vis[k];
            }
        }
        if (p.cfl_enable && solver != VTP_DFSPH)
            throw std::runtime_error(dom + " Extensions.cflEnable exists for DFSPH only (the other solvers step "
                                           "with the fixed timeStep)");
        if (p.particle_radius <= 0.f || p.time_step <= 0.f)
            throw std::runtime_error(dom + " particleRadius and timeStep must be positive");
        if (p.max_neighbor_num < 1) throw std::runtime_error(dom + " maxNeighborNum must be >= 1");
        if ((solver == VTP_IMM || solver == VTP_IMMCT) &&
            (p.polymer_phase < 0 || uint32_t(p.polymer_phase) >= p.phase_num) && solver == VTP_IMMCT)
            throw std::runtime_error(dom + " polymerPhase out of range");
        *out = p;
        return VTP_OK;
    } catch (const std::exception &e) {
        set_err(err, err_len, e.what());
        return VTP_EINVAL;
    }
}

extern "C" vtp_status vtp_params_default(int32_t solver, vtp_params *out) {
    if (solver < 0 || solver > 3) return VTP_EINVAL;
    return vtp_params_from_json(solver, kTemplate[solver], out, nullptr, 0);
}

namespace {
// Object-component templates
// (ConfigTemplates/<Solver>/<Solver>ObjectComponentConfigTemplate.json).
const char *kObjectTemplate[4] = {
    R"({"solverType":0,"exportFlag":false,"velocityStart":[0,0,0],"colorStart":[0,1,0]})",
    R"({"solverType":1,"exportFlag":false,"velocityStart":[0,0,0],"colorStart":[0,1,0]})",
    R"({"solverType":2,"exportFlag":false,"velocityStart":[0,0,0],"phaseFraction":[0.1,0.1]})",
    R"({"solverType":3,"exportFlag":false,"velocityStart":[0,0,0],"phaseFraction":[0.5,0.5]})",
};
}  // namespace

extern "C" const char *vtp_config_template(int32_t solver) {
    return (solver >= 0 && solver < 4) ? kTemplate[solver] : nullptr;
}

extern "C" const char *vtp_object_config_template(int32_t solver) {
    return (solver >= 0 && solver < 4) ? kObjectTemplate[solver] : nullptr;
}
