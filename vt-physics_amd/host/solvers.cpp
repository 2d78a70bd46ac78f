// solvers.cpp - the reference's Solver classes, SolverManager and Simulator
// over the C ABI (include/vtp.h).
//
// Behaviour follows Simulator/Runtime/Source/Solvers/PBF/PBFSolver.cpp (and
// its DFSPH / IMM / IMM-CT siblings): setConfig stores the document and checks
// its keys; attachObject validates the object and appends its particles in
// attach order, recording the export ranges in EXPORT.SolverRequired;
// initialize builds the device state; run ticks until animationTime and
// exports a frame whenever sim time passes k/exportFps; a failed tick makes
// the solver crashed for good.  The device work of one tick is one vtp_step.
#include <cstring>
#include <set>

#include "vt_physics.hpp"

namespace VT_Physics {

Solver::~Solver() = default;

namespace {

const char *kSolverName[4] = {"PBFSolver", "DFSPHSolver", "IMMSolver", "IMMCTSolver"};
const char *kDomain[4] = {"PBF", "DFSPH", "IMM", "IMMCT"};

// <Solver>ObjectComponentConfigRequiredKeys (PBFSolver.hpp:36-41 and siblings)
const std::vector<std::string> kObjectKeys[4] = {
    {"solverType", "exportFlag", "velocityStart", "colorStart"},
    {"solverType", "exportFlag", "velocityStart", "colorStart"},
    {"solverType", "exportFlag", "velocityStart", "phaseFraction"},
    {"solverType", "exportFlag", "velocityStart", "phaseFraction"},
};

bool is_phase_solver(SolverType t) { return t == IMM || t == IMMCT; }

}  // namespace

ParticleSolver::ParticleSolver(SolverType type, uint32_t /*cudaThreadSize: block sizes are per kernel here*/)
    : m_type(type) {
    m_configData = JsonHandler::loadSolverConfigTemplateJson(type);
    vtp_params_default(int32_t(type), &m_params);
    VTP_LOG_INFO(std::string(kSolverName[type]) + " Created.");
}

ParticleSolver::~ParticleSolver() {
    if (m_ctx) vtp_destroy(m_ctx);
}

json ParticleSolver::getSolverConfigTemplate() const { return JsonHandler::loadSolverConfigTemplateJson(m_type); }

json ParticleSolver::getSolverObjectComponentConfigTemplate() {
    return JsonHandler::loadSolverObjectComponentConfigTemplateJson(m_type);
}

bool ParticleSolver::checkConfig() const {
    vtp_params p;
    char err[512] = {0};
    if (vtp_params_from_json(int32_t(m_type), m_configData.dump().c_str(), &p, err, sizeof err) != VTP_OK) {
        VTP_LOG_ERROR(err);
        return false;
    }
    return true;
}

bool ParticleSolver::setConfig(json config) {
    m_configData = std::move(config);
    if (m_configData.empty()) {
        VTP_LOG_ERROR(std::string(kDomain[m_type]) + " config json is empty.");
        return false;
    }
    if (!checkConfig()) {
        VTP_LOG_ERROR(std::string("Invalid ") + kDomain[m_type] + " config file.");
        return false;
    }
    vtp_params_from_json(int32_t(m_type), m_configData.dump().c_str(), &m_params, nullptr, 0);
    if (is_phase_solver(m_type)) m_configData["EXPORT"]["Common"]["exportPhaseNum"] = m_params.phase_num;
    m_isConfigured = true;
    VTP_LOG_INFO(std::string(kSolverName[m_type]) + " Configured.");
    return true;
}

bool ParticleSolver::setConfigByFile(std::string config_file) {
    setConfig(JsonHandler::loadJson(config_file));
    return true;  // the reference ignores the outcome (PBFSolver.cpp:90-93)
}

bool ParticleSolver::attachObject(Object *obj) {
    const std::string who = obj ? "Object: " + std::to_string(obj->getID()) : std::string("Object: null");
    if (!obj || !obj->getObjectComponent()) {
        VTP_LOG_ERROR(who + " has no geometry component.");
        return false;
    }
    if (uint8_t(obj->getObjectComponent()->getType()) > 20) {
        VTP_LOG_ERROR(std::string(kSolverName[m_type]) + " only support particle geometry yet.");
        return false;
    }
    json &oc = obj->getObjectComponentConfig();
    const int mat = oc.contains("epmMaterial") && oc["epmMaterial"].is_number() ? int(oc["epmMaterial"].number()) : -1;
    if (mat != EPM_FLUID && mat != EPM_BOUNDARY) {
        VTP_LOG_ERROR(std::string(kSolverName[m_type]) + " unsupported material.");
        return false;
    }
    json &sc = obj->getSolverObjectComponentConfig();
    for (const auto &k : kObjectKeys[m_type])
        if (!sc.contains(k)) {
            VTP_LOG_ERROR(who + " has no " + kDomain[m_type] + " object component config key: " + k);
            return false;
        }
    if (int(sc["solverType"].number()) != int(m_type)) {
        VTP_LOG_ERROR(who + " has no " + kDomain[m_type] + " object component config.");
        return false;
    }
    std::vector<float> frac;
    if (is_phase_solver(m_type)) {
        frac = sc["phaseFraction"].floats();
        if (frac.size() != size_t(m_params.phase_num)) {
            VTP_LOG_ERROR(who + " phaseFraction size not match with solver.");
            return false;
        }
    }
    json &sr = m_configData["EXPORT"]["SolverRequired"];
    if (!sr.contains("exportObjectStartIndex")) {
        sr["exportObjectMaterials"] = json::array();
        sr["exportObjectStartIndex"] = json::array();
        sr["exportObjectEndIndex"] = json::array();
        sr["exportFlags"] = json::array();
    }
    sr["exportObjectMaterials"].push_back(mat);
    sr["exportObjectStartIndex"].push_back((unsigned long long)m_host_pos.size());
    sr["exportFlags"].push_back(sc["exportFlag"].boolean());

    const std::vector<float> el = obj->getObjectComponent()->getElements();
    const size_t n = el.size() / 3;
    const auto v0 = sc["velocityStart"].floats();
    float3 color{0.f, 0.f, 0.f};
    if (is_phase_solver(m_type)) {
        // colour = sum_k frac_k * phaseRestColor_k (IMMSolver.cpp:218-226)
        for (uint32_t k = 0; k < m_params.phase_num; ++k) {
            color.x += frac[k] * m_params.phase_rest_color[k][0];
            color.y += frac[k] * m_params.phase_rest_color[k][1];
            color.z += frac[k] * m_params.phase_rest_color[k][2];
        }
    } else {
        const auto c = sc["colorStart"].floats();
        color = {c[0], c[1], c[2]};
    }
    m_obj_start.push_back(uint32_t(m_host_pos.size()));
    m_obj_count.push_back(uint32_t(n));
    for (size_t i = 0; i < n; ++i) {
        m_host_pos.push_back({el[3 * i], el[3 * i + 1], el[3 * i + 2]});
        m_host_vel.push_back({v0[0], v0[1], v0[2]});
        m_host_mat.push_back(mat);
        m_host_color.push_back(color);
        for (float a : frac) m_host_phaseFrac.push_back(a);
    }
    sr["exportObjectEndIndex"].push_back((unsigned long long)m_host_pos.size());
    m_attached_objs.push_back(obj);
    return true;
}

bool ParticleSolver::attachObjects(std::vector<Object *> objs) {
    for (Object *o : objs)
        if (!attachObject(o)) return false;
    return true;
}

bool ParticleSolver::initialize() {
    if (m_isInitialized) return true;
    if (!m_configData.contains("EXPORT") || !ExportUtil::checkConfig(m_configData["EXPORT"])) {
        VTP_LOG_ERROR(std::string(kSolverName[m_type]) + " export config not available.");
        return false;
    }
    if (m_host_pos.empty()) {
        VTP_LOG_ERROR(std::string(kSolverName[m_type]) + " has no particles attached.");
        return false;
    }
    if (vtp_create_with_params(&m_ctx, &m_params) != VTP_OK) {
        VTP_LOG_ERROR(std::string(kSolverName[m_type]) + " fail to create device state: " + vtp_last_error(nullptr));
        m_ctx = nullptr;
        return false;
    }
    const size_t P = is_phase_solver(m_type) ? size_t(m_params.phase_num) : 0;
    for (size_t o = 0; o < m_obj_start.size(); ++o) {
        const uint32_t s = m_obj_start[o], n = m_obj_count[o];
        if (n == 0) continue;
        const vtp_status st = vtp_add_particles(
            m_ctx, 0, m_host_mat[s], &m_host_pos[s].x, &m_host_vel[s].x,
            P ? &m_host_phaseFrac[size_t(s) * P] : nullptr, n);
        if (st != VTP_OK) {
            VTP_LOG_ERROR(std::string(kSolverName[m_type]) + " fail to attach particles: " + vtp_last_error(m_ctx));
            return false;
        }
    }
    if (vtp_initialize(m_ctx) != VTP_OK) {
        VTP_LOG_ERROR(std::string(kSolverName[m_type]) + " fail to initialize: " + vtp_last_error(m_ctx));
        return false;
    }
    m_isInitialized = true;
    return true;
}

bool ParticleSolver::tick() {
    const float export_gap = 1.0f / float(m_configData["EXPORT"]["SolverRequired"]["exportFps"].number());
    const vtp_status st = vtp_step(m_ctx, 1);
    vtp_step_stats s;
    if (vtp_stats(m_ctx, &s) == VTP_OK) m_curSimTime = s.sim_time;
    m_frameCount++;
    if (m_curSimTime >= export_gap * float(m_outputFrameCount)) {
        m_outputFrameCount++;
        m_doExportFlag = true;
    }
    if (st != VTP_OK) {
        VTP_LOG_ERROR(std::string(kSolverName[m_type]) + " tick failed: " + vtp_last_error(m_ctx));
        m_isCrashed = true;
        return false;
    }
    return true;
}

bool ParticleSolver::run() {
    if (!m_isInitialized && !initialize()) return false;
    const float end = m_params.animation_time;
    const bool exporting = m_configData["EXPORT"]["SolverRequired"]["enable"].boolean();
    if (!exporting && !m_isCrashed) {
        // Nothing to write between ticks: let the library run the loop (graph
        // replays, no host sync per tick).  Frame counters are kept in step.
        const vtp_status st = vtp_run(m_ctx);
        vtp_step_stats s;
        if (vtp_stats(m_ctx, &s) == VTP_OK) {
            m_curSimTime = s.sim_time;
            m_frameCount = uint32_t(s.steps);
        }
        if (st != VTP_OK) {
            VTP_LOG_ERROR(std::string(kSolverName[m_type]) + " is crashed: " + vtp_last_error(m_ctx));
            m_isCrashed = true;
            return false;
        }
        return true;
    }
    while (m_curSimTime < end) {
        if (m_isCrashed) {
            VTP_LOG_ERROR(std::string(kSolverName[m_type]) + " is crashed.");
            return false;
        }
        tick();
        exportData();
        VTP_LOG_INFO(std::string(kSolverName[m_type]) + " current frame: " + std::to_string(m_frameCount));
    }
    return true;
}

bool ParticleSolver::tickNsteps(uint32_t n) {
    if (!m_isInitialized && !initialize()) return false;
    for (uint32_t i = 0; i < n; ++i) {
        if (m_isCrashed) {
            VTP_LOG_ERROR(std::string(kSolverName[m_type]) + " is crashed.");
            return false;
        }
        tick();
        exportData();
        VTP_LOG_INFO(std::string(kSolverName[m_type]) + " current frame: " + std::to_string(m_frameCount));
    }
    return true;
}

bool ParticleSolver::readField(int32_t field, std::vector<float> &out) const {
    if (!m_ctx) return false;
    const size_t n = m_host_pos.size();
    size_t comps = 1;
    if (field == VTP_FIELD_POS || field == VTP_FIELD_VEL || field == VTP_FIELD_COLOR ||
        field == VTP_FIELD_ACC || field == VTP_FIELD_SURFACE_NORMAL || field == VTP_FIELD_VEL_ADV)
        comps = 3;
    else if (field == VTP_FIELD_VOL_FRAC)
        comps = size_t(m_params.phase_num);
    else if (field == VTP_FIELD_VEL_PHASE || field == VTP_FIELD_VEL_DRIFT)
        comps = 3 * size_t(m_params.phase_num);
    else if (field == VTP_FIELD_Q || field == VTP_FIELD_VEL_GRAD || field == VTP_FIELD_VISE_STRESS)
        comps = 9;
    out.resize(n * comps);
    return vtp_read(m_ctx, field, out.data(), out.size() * sizeof(float)) == VTP_OK;
}

// PBFSolver.cpp:248-324 / IMMSolver.cpp exportData.  MERGE: one file per
// material that has an exporting object, holding every object of that
// material, in <dir>/<Fluid|Boundary>/<frame>.ply (positions only for
// PBF/DFSPH; positions, colour and phase fractions for IMM/IMM-CT).  SPLIT:
// one file per exporting object in <dir>/<object name>/<prefix><frame>.ply
// with positions and colour.
void ParticleSolver::exportData() {
    json &ec = m_configData["EXPORT"];
    if (!ec["SolverRequired"]["enable"].boolean() || !m_doExportFlag) return;
    m_doExportFlag = false;
    const json &sr = ec["SolverRequired"];
    const std::string policy = sr["exportGroupPolicy"].str();
    if (!supportedExportPolicy.count(policy)) {
        VTP_LOG_ERROR(std::string(kSolverName[m_type]) + " exportGroupPolicy" + policy + " not supported.");
        return;
    }
    std::vector<float> pos, color, frac;
    if (!readField(VTP_FIELD_POS, pos)) {
        VTP_LOG_ERROR(std::string(kSolverName[m_type]) + " export read failed: " + vtp_last_error(m_ctx));
        return;
    }
    const bool phases = is_phase_solver(m_type);
    const size_t P = phases ? size_t(m_params.phase_num) : 0;
    if (phases && (!readField(VTP_FIELD_COLOR, color) || !readField(VTP_FIELD_VOL_FRAC, frac))) {
        VTP_LOG_ERROR(std::string(kSolverName[m_type]) + " export read failed: " + vtp_last_error(m_ctx));
        return;
    }
    auto pos_of = [&](size_t i) { return float3{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]}; };
    auto color_of = [&](size_t i) {
        return phases ? float3{color[3 * i], color[3 * i + 1], color[3 * i + 2]} : m_host_color[i];
    };
    const size_t nobj = sr["exportObjectMaterials"].size();
    if (policy == "MERGE") {
        std::set<int> done;
        for (size_t i = 0; i < nobj; ++i) {
            if (!sr["exportFlags"][i].boolean()) continue;
            const int mat = int(sr["exportObjectMaterials"][i].number());
            if (!done.insert(mat).second) continue;
            std::vector<float3> p, c;
            std::vector<float> ph;
            for (size_t j = 0; j < nobj; ++j) {
                if (int(sr["exportObjectMaterials"][j].number()) != mat) continue;
                const size_t s = size_t(sr["exportObjectStartIndex"][j].number());
                const size_t e = size_t(sr["exportObjectEndIndex"][j].number());
                for (size_t k = s; k < e; ++k) {
                    p.push_back(pos_of(k));
                    if (phases) {
                        c.push_back(color_of(k));
                        for (size_t q = 0; q < P; ++q) ph.push_back(frac[k * P + q]);
                    }
                }
            }
            json tmp = ec;
            tmp["Common"]["exportTargetDir"] = ec["Common"]["exportTargetDir"].str() + "/" + EPMString.at(mat);
            tmp["Common"]["exportFilePrefix"] = std::to_string(m_outputFrameCount);
            ExportUtil::exportData(tmp, p, c, ph);
        }
    } else {
        for (size_t i = 0; i < nobj; ++i) {
            if (!sr["exportFlags"][i].boolean()) continue;
            const size_t s = size_t(sr["exportObjectStartIndex"][i].number());
            const size_t e = size_t(sr["exportObjectEndIndex"][i].number());
            std::vector<float3> p, c;
            for (size_t k = s; k < e; ++k) {
                p.push_back(pos_of(k));
                c.push_back(color_of(k));
            }
            json tmp = ec;
            tmp["Common"]["exportTargetDir"] =
                ec["Common"]["exportTargetDir"].str() + "/" + m_attached_objs[i]->getName();
            tmp["Common"]["exportFilePrefix"] =
                ec["Common"]["exportFilePrefix"].str() + std::to_string(m_outputFrameCount);
            ExportUtil::exportData(tmp, p, c);
        }
    }
}

bool ParticleSolver::reset() {
    m_configData = JsonHandler::loadSolverConfigTemplateJson(m_type);
    return true;
}

void ParticleSolver::destroy() {
    if (m_ctx) vtp_destroy(m_ctx);
    m_ctx = nullptr;
    m_isInitialized = false;
    m_attached_objs.clear();
    VTP_LOG_INFO(std::string(kSolverName[m_type]) + " Destroyed.");
}

// ---------------------------------------------------------------- managers
Solver *SolverManager::createSolver(SolverType solverType) {
    Solver *s = nullptr;
    switch (solverType) {
        case PBF: s = new pbf::PBFSolver(m_cuda_threadNum_per_block); break;
        case DFSPH: s = new dfsph::DFSPHSolver(m_cuda_threadNum_per_block); break;
        case IMM: s = new imm::IMMSolver(m_cuda_threadNum_per_block); break;
        case IMMCT: s = new immct::IMMCTSolver(m_cuda_threadNum_per_block); break;
        default: return nullptr;
    }
    m_solvers.push_back(s);
    return s;
}

void SolverManager::clear() {
    for (Solver *s : m_solvers) {
        s->destroy();
        delete s;
    }
    m_solvers.clear();
}

Simulator &Simulator::getInstance() {
    static Simulator instance;
    return instance;
}
ObjectManager *Simulator::getObjectManager() { return &m_objectManager; }
SolverManager *Simulator::getSolverManager() { return &m_solverManager; }
void Simulator::terminate() {
    m_solverManager.clear();
    m_objectManager.clear();
}
void Simulator::clean() {}  // no static component templates to free here

}  // namespace VT_Physics
