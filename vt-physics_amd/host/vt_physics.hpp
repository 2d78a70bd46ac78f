// vt_physics.hpp - the reference's host-side C++ API over the C ABI.
//
// Same class names, method names, JSON keys, error behaviour (bool + a log
// line, sticky crash flag) and attach-order conventions as the reference's
// runtime, so a program written against it compiles against this header with
// only the include changed:
//   Simulator/Runtime/Include/Framework/Solver.hpp:20-62        Solver, SolverType
//   Simulator/Runtime/Include/Framework/Object.hpp:13-42        Object
//   Simulator/Runtime/Include/Framework/ObjectComponents.hpp    ObjectType, components
//   Simulator/Runtime/Include/Framework/Material.hpp:14-22      EPM_FLUID / EPM_BOUNDARY
//   Simulator/Runtime/Include/Manager/{Simulator,ObjectManager,SolverManager}.hpp
//   Simulator/Library/Include/Model/{ModelHandler,ExportUtil}.hpp
//   Simulator/Runtime/Include/Solvers/{PBF,DFSPH,IMM,IMM-CT}/*Solver.hpp
//
// Every solver is a thin owner of one vtp_ctx (include/vtp.h): the particle
// state lives on the GPU in the library's own layout and is read back in
// attach order only for export.  `json` is vtp::Json (a small self-contained
// JSON value with the subset of nlohmann::json's interface used here).
#ifndef VT_PHYSICS_HOST_HPP
#define VT_PHYSICS_HOST_HPP

#include <climits>
#include <cstdint>
#include <set>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/vtp.h"
#include "json.hpp"

namespace VT_Physics {

using json = vtp::Json;

struct float3 {
    float x, y, z;
};

// ---- logging (the reference logs through spdlog; here stderr, filtered by
// the VTP_LOG environment variable: 0 errors, 1 +warnings (default), 2 +info)
void log_message(int level, const std::string &msg);
#define VTP_LOG_ERROR(m) ::VT_Physics::log_message(0, (m))
#define VTP_LOG_WARNING(m) ::VT_Physics::log_message(1, (m))
#define VTP_LOG_INFO(m) ::VT_Physics::log_message(2, (m))

// ---- materials (Material.hpp:14-22)
typedef enum eParticleMaterial : int {
    epm_Fluid = 0,
    epm_Boundary = 1,
    epm_Null = INT_MAX
} ParticleMaterial;
#define EPM_FLUID static_cast<int>(::VT_Physics::ParticleMaterial::epm_Fluid)
#define EPM_BOUNDARY static_cast<int>(::VT_Physics::ParticleMaterial::epm_Boundary)
inline const std::unordered_map<int, std::string> EPMString = {{EPM_FLUID, "Fluid"},
                                                               {EPM_BOUNDARY, "Boundary"}};

// ---- object types (ObjectComponents.hpp:18-32)
typedef enum eObjectType : uint8_t {
    Particle_Common = 0,
    Particle_Sphere = 1,
    Particle_Cube = 2,
    Particle_Cylinder = 3,
    Particle_Plane = 4,
    Particle_Box = 5,
    MeshGeometry = 20,
    Mesh_Sphere = 21,
    Mesh_Cube = 22,
    Mesh_Cylinder = 23,
    Mesh_Plane = 24,
    Mesh_Particle = 25,
    OBJ_NULL_TYPE = UINT8_MAX
} ObjectType;

// ---- solver types (Solver.hpp:20-25)
typedef enum eSolverType : uint8_t { PBF = 0, DFSPH = 1, IMM = 2, IMMCT = 3 } SolverType;

// ---- geometry generation and file I/O (ModelHandler.cpp, ExportUtil.cpp)
class ModelHandler {
public:
    // Dispatch on config["genType"] (ModelHandler.cpp:13-41).
    static std::vector<float3> generateObjectElements(const json &config);
    // Dispatch on config["loadType"]; particle files by extension (:176-229).
    static std::vector<float3> loadObjectElements(const json &config);
    static std::vector<float3> generateParticleCubeElements(const json &config);
    static std::vector<float3> generateParticleCylinderElements(const json &config);
    static std::vector<float3> generateParticleSphereElements(const json &config);
    static std::vector<float3> generateParticleBoxElements(const json &config);
    static std::vector<float3> generateParticlePlaneElements(const json &config);
    static std::vector<float3> loadParticleElements(const json &config);
    // Vertex positions of a PLY point cloud (ascii or binary, either endian).
    static std::vector<float3> load_PLY_ParticleElements(const std::string &filePath);
    static std::string getFileExtensionUpper(const std::string &filePath);
};

inline const std::vector<std::string> exportConfigRequiredCommonKeys = {"exportTargetDir", "exportFilePrefix",
                                                                        "exportFileType"};
inline const std::vector<std::string> exportConfigRequiredSolverRequiredKeys = {
    "enable", "exportFps", "exportGroupPolicy", "exportObjectMaterials", "exportObjectStartIndex",
    "exportObjectEndIndex", "exportFlags"};
inline const std::set<std::string> supportedFileType = {"PLY"};
inline const std::set<std::string> supportedExportPolicy = {"MERGE", "SPLIT"};

class ExportUtil {
public:
    static void exportData(const json &exportConfig, const std::vector<float3> &pos,
                           const std::vector<float3> &color = {}, const std::vector<float> &phase = {});
    static void exportData(const json &exportConfig, std::vector<float> &pos);
    static void exportAsPly(const json &exportConfig, const std::vector<float3> &pos,
                            const std::vector<float3> &color, const std::vector<float> &phase);
    static bool checkConfig(const json &config);
};

class JsonHandler {
public:
    static json loadJson(const std::string &path);
    static bool saveJson(const std::string &path, const json &j);
    static json loadSolverConfigTemplateJson(SolverType t);
    static json loadSolverObjectComponentConfigTemplateJson(SolverType t);
    static json loadPBFConfigTemplateJson() { return loadSolverConfigTemplateJson(PBF); }
    static json loadDFSPHConfigTemplateJson() { return loadSolverConfigTemplateJson(DFSPH); }
    static json loadIMMConfigTemplateJson() { return loadSolverConfigTemplateJson(IMM); }
    static json loadIMMCTConfigTemplateJson() { return loadSolverConfigTemplateJson(IMMCT); }
};

// ---- object components (ObjectComponents.hpp:34-199)
struct ObjectTypeComponent {
    virtual ~ObjectTypeComponent() = default;
    virtual ObjectType getType() const = 0;
    virtual ObjectTypeComponent *clone() const = 0;
    virtual json getConfig() = 0;
    virtual void update(json &config) = 0;
    virtual std::vector<float> getElements() = 0;
};

// One implementation for every particle geometry: the type decides the
// default config (the reference's per-struct member defaults), the keys that
// must be present, and the generator (or file loader) that fills `pos`.
struct ParticleGeometryComponent : public ObjectTypeComponent {
    explicit ParticleGeometryComponent(ObjectType type = Particle_Common) : m_type(type) {}
    std::vector<float3> pos;
    float particleRadius{0.05f};
    uint8_t epmMaterial{UINT8_MAX};
    ObjectType getType() const override { return m_type; }
    ObjectTypeComponent *clone() const override { return new ParticleGeometryComponent(*this); }
    json getConfig() override;
    void update(json &config) override;
    std::vector<float> getElements() override;

private:
    bool configIsComplete(const json &config) const;
    ObjectType m_type;
};

struct MeshGeometryComponent : public ObjectTypeComponent {
    ObjectType getType() const override { return MeshGeometry; }
    ObjectTypeComponent *clone() const override { return new MeshGeometryComponent(*this); }
    json getConfig() override { return json::object(); }
    void update(json &) override {}
    std::vector<float> getElements() override { return {}; }
};

// ---- Object (Object.hpp:13-42; Object.cpp)
class Object {
public:
    explicit Object(long long id);
    ~Object();
    Object(const Object &) = delete;
    Object &operator=(const Object &) = delete;
    void transferTo(ObjectType type);
    json &getObjectComponentConfig();
    void attachSpecificSolverObjectComponentConfig(json specificSolverObjectComponentConfig);
    json &getSolverObjectComponentConfig();
    ObjectTypeComponent *getObjectComponent();
    long long getID() const;
    void rename(std::string newName);
    std::string getName() const;
    bool update();
    void reset();

private:
    long long m_id;
    std::string m_name;
    ObjectTypeComponent *m_objectTypeComponent{nullptr};
    json m_objectComponentConfig;
    json m_solverSpecificComponentConfig;
};

// Owns the objects it creates; clear() (or Simulator::terminate) frees them.
// Like the reference, copies share the pointers and free nothing on their own.
class ObjectManager {
public:
    // ObjectManager.cpp:22-41: one object per entry of allConfiguration["OBJECTS"].
    std::vector<Object *> createObjectsFromJson(json allConfiguration);
    Object *createObject(ObjectType type = OBJ_NULL_TYPE);
    void clear();

private:
    long long m_objectNum{0};
    std::vector<Object *> m_objects;
};

// ---- Solver (Solver.hpp:27-62)
class Solver {
public:
    Solver() = default;
    virtual ~Solver() = 0;
    virtual json getSolverConfigTemplate() const = 0;
    virtual bool setConfig(json config) = 0;
    virtual bool setConfigByFile(std::string config_file) = 0;
    virtual json getSolverObjectComponentConfigTemplate() = 0;
    virtual bool run() = 0;
    virtual bool tickNsteps(uint32_t n) = 0;
    virtual bool attachObject(Object *obj) = 0;
    virtual bool attachObjects(std::vector<Object *> objs) = 0;
    virtual bool initialize() = 0;
    virtual bool reset() = 0;
    virtual void destroy() = 0;

protected:
    virtual bool tick() = 0;
    virtual bool checkConfig() const = 0;
};

// Shared body of the four solvers: each reference solver class differs only
// in its config domain, the per-object keys it requires, and what it exports
// (PBFSolver.cpp, DFSPHSolver.cpp, IMMSolver.cpp, IMMCTSolver.cpp).
class ParticleSolver : public Solver {
public:
    ~ParticleSolver() override;
    json getSolverConfigTemplate() const override;
    bool setConfig(json config) override;
    bool setConfigByFile(std::string config_file) override;
    json getSolverObjectComponentConfigTemplate() override;
    bool run() override;
    bool tickNsteps(uint32_t n) override;
    bool attachObject(Object *obj) override;
    bool attachObjects(std::vector<Object *> objs) override;
    bool initialize() override;
    bool reset() override;
    void destroy() override;

    // Not in the reference: access to the device context for callers that
    // want fields without going through export.
    vtp_ctx *context() const { return m_ctx; }
    SolverType solverType() const { return m_type; }
    uint32_t particleNum() const { return uint32_t(m_host_pos.size()); }
    // Copies one field (attach order) into `out`; false on error.
    bool readField(int32_t field, std::vector<float> &out) const;

protected:
    ParticleSolver(SolverType type, uint32_t cudaThreadSize);
    bool tick() override;
    bool checkConfig() const override;
    void exportData();

    SolverType m_type;
    json m_configData;
    vtp_params m_params{};
    vtp_ctx *m_ctx{nullptr};
    bool m_isConfigured{false};
    bool m_isInitialized{false};
    bool m_isCrashed{false};
    uint32_t m_frameCount{0};
    uint32_t m_outputFrameCount{0};
    bool m_doExportFlag{false};
    float m_curSimTime{0.f};
    std::vector<Object *> m_attached_objs;
    std::vector<float3> m_host_pos;
    std::vector<float3> m_host_vel;
    std::vector<int> m_host_mat;
    std::vector<float3> m_host_color;
    std::vector<float> m_host_phaseFrac;
    std::vector<uint32_t> m_obj_start;   // per attached object: first particle
    std::vector<uint32_t> m_obj_count;   // and particle count
};

namespace pbf {
class PBFSolver : public ParticleSolver {
public:
    PBFSolver() = delete;
    explicit PBFSolver(uint32_t cudaThreadSize) : ParticleSolver(SolverType::PBF, cudaThreadSize) {}
};
}  // namespace pbf
namespace dfsph {
class DFSPHSolver : public ParticleSolver {
public:
    DFSPHSolver() = delete;
    explicit DFSPHSolver(uint32_t cudaThreadSize) : ParticleSolver(SolverType::DFSPH, cudaThreadSize) {}
};
}  // namespace dfsph
namespace imm {
class IMMSolver : public ParticleSolver {
public:
    IMMSolver() = delete;
    explicit IMMSolver(uint32_t cudaThreadSize) : ParticleSolver(SolverType::IMM, cudaThreadSize) {}
};
}  // namespace imm
namespace immct {
class IMMCTSolver : public ParticleSolver {
public:
    IMMCTSolver() = delete;
    explicit IMMCTSolver(uint32_t cudaThreadSize) : ParticleSolver(SolverType::IMMCT, cudaThreadSize) {}
};
}  // namespace immct

// ---- managers (SolverManager.cpp:19-48; Simulator.hpp)
class SolverManager {
public:
    SolverManager() = default;
    Solver *createSolver(SolverType solverType);
    void clear();

private:
    uint32_t m_cuda_threadNum_per_block{256};
    std::vector<Solver *> m_solvers;
};

class Simulator final {
public:
    static Simulator &getInstance();
    ObjectManager *getObjectManager();
    SolverManager *getSolverManager();
    void terminate();
    static void clean();

private:
    ObjectManager m_objectManager;
    SolverManager m_solverManager;
};

}  // namespace VT_Physics

#endif  // VT_PHYSICS_HOST_HPP
