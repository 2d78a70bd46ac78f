// objects.cpp - objects, generators, PLY import/export and JSON files of the
// host API (vt_physics.hpp).
//
// Generators follow Simulator/Library/Source/Model/ModelHandler.cpp:43-154
// step for step in FP32 (the loops accumulate `x += diameter` in float, the
// box evaluates `lb + r + 2.0*r*i` in double and rounds once), so the particle
// counts and positions equal what the reference attaches.  Compile with
// -ffp-contract=off: a fused multiply-add would change the sphere test.
#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

#include "vt_physics.hpp"

namespace VT_Physics {

// ---------------------------------------------------------------- logging
void log_message(int level, const std::string &msg) {
    static const int threshold = [] {
        const char *e = std::getenv("VTP_LOG");
        return e ? std::atoi(e) : 1;
    }();
    if (level > threshold) return;
    static const char *tag[3] = {"error", "warning", "info"};
    std::fprintf(stderr, "[vt-physics %s] %s\n", tag[level < 0 ? 0 : level > 2 ? 2 : level], msg.c_str());
}

namespace {

float f(const json &j) { return float(j.number()); }

std::vector<float> fvec(const json &j) { return j.floats(); }

bool make_dirs(const std::string &path) {
    if (path.empty()) return true;
    std::string cur;
    size_t i = 0;
    if (path[0] == '/') {
        cur = "/";
        i = 1;
    }
    while (i <= path.size()) {
        size_t j = path.find('/', i);
        if (j == std::string::npos) j = path.size();
        cur += path.substr(i, j - i);
        if (!cur.empty() && cur != "/") {
            struct stat st;
            if (stat(cur.c_str(), &st) != 0 && mkdir(cur.c_str(), 0755) != 0) return false;
        }
        cur += '/';
        i = j + 1;
    }
    return true;
}

}  // namespace

// ---------------------------------------------------------------- JsonHandler
json JsonHandler::loadJson(const std::string &path) {
    std::ifstream in(path);
    if (!in) {
        VTP_LOG_ERROR("JsonHandler: cannot open " + path);
        return json();
    }
    std::stringstream ss;
    ss << in.rdbuf();
    try {
        return json::parse(ss.str());
    } catch (const std::exception &e) {
        VTP_LOG_ERROR("JsonHandler: " + path + ": " + e.what());
        return json();
    }
}

bool JsonHandler::saveJson(const std::string &path, const json &j) {
    std::ofstream out(path);
    if (!out) return false;
    out << j.dump(2);
    return bool(out);
}

json JsonHandler::loadSolverConfigTemplateJson(SolverType t) {
    const char *txt = vtp_config_template(int32_t(t));
    return txt ? json::parse(txt) : json();
}

json JsonHandler::loadSolverObjectComponentConfigTemplateJson(SolverType t) {
    const char *txt = vtp_object_config_template(int32_t(t));
    return txt ? json::parse(txt) : json();
}

// ---------------------------------------------------------------- ModelHandler
std::vector<float3> ModelHandler::generateObjectElements(const json &config) {
    if (!config.contains("genType")) {
        VTP_LOG_ERROR("No generation type specified.");
        return {};
    }
    switch (int(config["genType"].number())) {
        case 1: return generateParticleSphereElements(config);
        case 2: return generateParticleCubeElements(config);
        case 3: return generateParticleCylinderElements(config);
        case 4: return generateParticlePlaneElements(config);
        case 5: return generateParticleBoxElements(config);
        case 0:    // particle files go through loadObjectElements
        case 20:   // meshes are not generated
            return {};
        default:
            VTP_LOG_ERROR("Unknown or Unsupported generation type specified.");
            return {};
    }
}

// z outer, y, x inner; each coordinate accumulated in float (:43-67).
std::vector<float3> ModelHandler::generateParticleCubeElements(const json &config) {
    std::vector<float3> out;
    const float r = f(config["particleRadius"]);
    const auto lb = fvec(config["lb"]);
    const auto size = fvec(config["size"]);
    const float d = 2 * r;
    for (float z = r + lb[2]; z < lb[2] + size[2]; z += d)
        for (float y = r + lb[1]; y < lb[1] + size[1]; y += d)
            for (float x = r + lb[0]; x < lb[0] + size[0]; x += d) out.push_back({x, y, z});
    return out;
}

// Layers top-down from the top disc; per x a chord of half-length
// R*sqrt(1-cos^2) in z (:69-92).
std::vector<float3> ModelHandler::generateParticleCylinderElements(const json &config) {
    std::vector<float3> out;
    const float r = f(config["particleRadius"]);
    const auto c = fvec(config["center"]);
    const float R = f(config["bottomAreaRadius"]);
    const float height = f(config["height"]);
    const float d = 2 * r;
    const float tx = c[0], ty = c[1] + height / 2, tz = c[2];
    for (float y = ty - r; y >= ty - height; y -= d) {
        for (float x = (tx - R) + r; x <= tx + R; x += d) {
            const float m_cos = std::fabs(tx - x) / R;
            const float len = R * std::sqrt(1.0f - m_cos * m_cos);
            for (float z = (tz - len) + r; z <= tz + len; z += d) out.push_back({x, y, z});
        }
    }
    return out;
}

// Lattice points of spacing 2r within the ball, i/j/k in [-n, n] with
// n = ceil(R/2r) (:94-119).
std::vector<float3> ModelHandler::generateParticleSphereElements(const json &config) {
    std::vector<float3> out;
    const float r = f(config["particleRadius"]);
    const auto c = fvec(config["volumeCenter"]);
    const float R = f(config["volumeRadius"]);
    const float gap = r * 2.0f;
    const int n = int(std::ceil(R / gap));
    for (int i = -n; i <= n; ++i)
        for (int j = -n; j <= n; ++j)
            for (int k = -n; k <= n; ++k) {
                const float3 p = {float(i) * gap + c[0], float(j) * gap + c[1], float(k) * gap + c[2]};
                const float dx = p.x - c[0], dy = p.y - c[1], dz = p.z - c[2];
                const float d2 = (dx * dx + dy * dy) + dz * dz;
                if (d2 <= R * R) out.push_back(p);
            }
    return out;
}

// (n+1)^3 lattice of spacing 2r, n = int((size + r)/(2r)) computed in double;
// a point is kept when it lies in one of the layerNum outer layers on the low
// side or layerNum+1 on the high side (the `>= n - layerNum` test) (:124-154).
std::vector<float3> ModelHandler::generateParticleBoxElements(const json &config) {
    std::vector<float3> out;
    const float r = f(config["particleRadius"]);
    const auto lb = fvec(config["lb"]);
    const auto size = fvec(config["size"]);
    const int L = int(config["layerNum"].number());
    int num[3];
    for (int a = 0; a < 3; ++a) num[a] = int(double(size[a] + r) / (2.0 * double(r)));
    for (int i = 0; i <= num[0]; ++i)
        for (int j = 0; j <= num[1]; ++j)
            for (int k = 0; k <= num[2]; ++k) {
                if (i < L || i >= num[0] - L || j < L || j >= num[1] - L || k < L || k >= num[2] - L) {
                    float3 p;
                    p.x = float(double(lb[0] + r) + 2.0 * double(r) * i);
                    p.y = float(double(lb[1] + r) + 2.0 * double(r) * j);
                    p.z = float(double(lb[2] + r) + 2.0 * double(r) * k);
                    out.push_back(p);
                }
            }
    return out;
}

// z outer, layerNum layers in y, x inner (:155-172).
std::vector<float3> ModelHandler::generateParticlePlaneElements(const json &config) {
    std::vector<float3> out;
    const float r = f(config["particleRadius"]);
    const auto lb = fvec(config["lb"]);
    const auto size = fvec(config["size"]);
    const int L = int(config["layerNum"].number());
    const float d = 2 * r;
    for (float z = r + lb[2]; z < lb[2] + size[2]; z += d) {
        float y = r + lb[1];
        for (int cnt = 0; cnt < L; ++cnt, y += d)
            for (float x = r + lb[0]; x < lb[0] + size[0]; x += d) out.push_back({x, y, z});
    }
    return out;
}

std::vector<float3> ModelHandler::loadObjectElements(const json &config) {
    if (config.contains("loadType") && config["loadType"].number() < 20) return loadParticleElements(config);
    return {};
}

std::string ModelHandler::getFileExtensionUpper(const std::string &filePath) {
    const size_t dot = filePath.find_last_of('.');
    if (dot != std::string::npos && dot != 0 && dot != filePath.size() - 1) {
        std::string ext = filePath.substr(dot + 1);
        std::transform(ext.begin(), ext.end(), ext.begin(), [](unsigned char ch) { return char(std::toupper(ch)); });
        return ext;
    }
    return "";
}

std::vector<float3> ModelHandler::loadParticleElements(const json &config) {
    const std::string path = config["particleGeometryPath"].str();
    const std::string ext = getFileExtensionUpper(path);
    if (supportedFileType.count(ext) && ext == "PLY") return load_PLY_ParticleElements(path);
    VTP_LOG_ERROR("Unsupported file type: " + ext);
    return {};
}

namespace {

struct PlyProp {
    std::string name;
    int type = 0;       // byte size of the scalar (or of list items)
    bool is_float = false;
    bool is_signed = false;
    bool is_list = false;
    int count_type = 0;  // list length scalar size
};

bool ply_type(const std::string &t, int &size, bool &fl, bool &sg) {
    static const struct {
        const char *n;
        int s;
        bool f, g;
    } tab[] = {{"char", 1, false, true},    {"int8", 1, false, true},    {"uchar", 1, false, false},
               {"uint8", 1, false, false},  {"short", 2, false, true},   {"int16", 2, false, true},
               {"ushort", 2, false, false}, {"uint16", 2, false, false}, {"int", 4, false, true},
               {"int32", 4, false, true},   {"uint", 4, false, false},   {"uint32", 4, false, false},
               {"float", 4, true, true},    {"float32", 4, true, true},  {"double", 8, true, true},
               {"float64", 8, true, true}};
    for (const auto &e : tab)
        if (t == e.n) {
            size = e.s;
            fl = e.f;
            sg = e.g;
            return true;
        }
    return false;
}

double ply_scalar(const unsigned char *p, int size, bool fl, bool sg, bool swap) {
    unsigned char b[8];
    for (int i = 0; i < size; ++i) b[i] = swap ? p[size - 1 - i] : p[i];
    if (fl) {
        if (size == 4) {
            float v;
            std::memcpy(&v, b, 4);
            return v;
        }
        double v;
        std::memcpy(&v, b, 8);
        return v;
    }
    switch (size) {
        case 1: return sg ? double(int8_t(b[0])) : double(b[0]);
        case 2: {
            uint16_t u;
            std::memcpy(&u, b, 2);
            return sg ? double(int16_t(u)) : double(u);
        }
        default: {
            uint32_t u;
            std::memcpy(&u, b, 4);
            return sg ? double(int32_t(u)) : double(u);
        }
    }
}

}  // namespace

// Vertex positions of the "vertex" element (x, y, z properties, any scalar
// type).  The reference reads PLY through assimp (ModelHandler.cpp:206-229),
// which for a point cloud yields exactly these vertices in file order.
std::vector<float3> ModelHandler::load_PLY_ParticleElements(const std::string &filePath) {
    std::ifstream in(filePath, std::ios::binary);
    if (!in) {
        VTP_LOG_ERROR("PLY: cannot open " + filePath);
        return {};
    }
    std::string line;
    std::getline(in, line);
    if (line.rfind("ply", 0) != 0) {
        VTP_LOG_ERROR("PLY: missing magic in " + filePath);
        return {};
    }
    enum { ASCII, BIN_LE, BIN_BE } fmt = ASCII;
    struct Elem {
        std::string name;
        size_t count;
        std::vector<PlyProp> props;
    };
    std::vector<Elem> elems;
    bool header_ok = false;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        std::istringstream ls(line);
        std::string kw;
        ls >> kw;
        if (kw == "format") {
            std::string f;
            ls >> f;
            fmt = f == "ascii" ? ASCII : f == "binary_little_endian" ? BIN_LE : BIN_BE;
        } else if (kw == "element") {
            Elem e;
            ls >> e.name >> e.count;
            elems.push_back(e);
        } else if (kw == "property" && !elems.empty()) {
            PlyProp p;
            std::string t;
            ls >> t;
            if (t == "list") {
                std::string ct, it;
                ls >> ct >> it >> p.name;
                bool f1, s1;
                if (!ply_type(ct, p.count_type, f1, s1) || !ply_type(it, p.type, p.is_float, p.is_signed)) {
                    VTP_LOG_ERROR("PLY: bad list type in " + filePath);
                    return {};
                }
                p.is_list = true;
            } else {
                ls >> p.name;
                if (!ply_type(t, p.type, p.is_float, p.is_signed)) {
                    VTP_LOG_ERROR("PLY: bad property type '" + t + "' in " + filePath);
                    return {};
                }
            }
            elems.back().props.push_back(p);
        } else if (kw == "end_header") {
            header_ok = true;
            break;
        }
    }
    if (!header_ok) {
        VTP_LOG_ERROR("PLY: no end_header in " + filePath);
        return {};
    }
    const bool swap = fmt == BIN_BE;  // this host is little-endian
    std::vector<float3> out;
    for (const auto &e : elems) {
        const bool is_vertex = e.name == "vertex";
        int ix = -1, iy = -1, iz = -1;
        for (size_t k = 0; k < e.props.size(); ++k) {
            if (e.props[k].name == "x") ix = int(k);
            if (e.props[k].name == "y") iy = int(k);
            if (e.props[k].name == "z") iz = int(k);
        }
        if (is_vertex && (ix < 0 || iy < 0 || iz < 0)) {
            VTP_LOG_ERROR("PLY: vertex element without x/y/z in " + filePath);
            return {};
        }
        if (is_vertex) out.reserve(e.count);
        std::vector<double> v(e.props.size());
        for (size_t n = 0; n < e.count; ++n) {
            for (size_t k = 0; k < e.props.size(); ++k) {
                const auto &p = e.props[k];
                if (fmt == ASCII) {
                    double val;
                    if (p.is_list) {
                        double cnt;
                        if (!(in >> cnt)) goto truncated;
                        for (long c = 0; c < long(cnt); ++c)
                            if (!(in >> val)) goto truncated;
                        v[k] = 0;
                    } else {
                        if (!(in >> val)) goto truncated;
                        v[k] = val;
                    }
                } else {
                    unsigned char buf[8];
                    if (p.is_list) {
                        if (!in.read(reinterpret_cast<char *>(buf), p.count_type)) goto truncated;
                        const long cnt = long(ply_scalar(buf, p.count_type, false, false, swap));
                        in.seekg(long(cnt) * p.type, std::ios::cur);
                        v[k] = 0;
                    } else {
                        if (!in.read(reinterpret_cast<char *>(buf), p.type)) goto truncated;
                        v[k] = ply_scalar(buf, p.type, p.is_float, p.is_signed, swap);
                    }
                }
            }
            if (is_vertex) out.push_back({float(v[ix]), float(v[iy]), float(v[iz])});
        }
        if (is_vertex) return out;
    }
    VTP_LOG_ERROR("PLY: no vertex element in " + filePath);
    return {};
truncated:
    VTP_LOG_ERROR("PLY: truncated data in " + filePath);
    return {};
}

// ---------------------------------------------------------------- ExportUtil
void ExportUtil::exportData(const json &exportConfig, const std::vector<float3> &pos,
                            const std::vector<float3> &color, const std::vector<float> &phase) {
    VTP_LOG_INFO("Exporting data...");
    const std::string type = exportConfig["Common"]["exportFileType"].str();
    if (!supportedFileType.count(type)) {
        VTP_LOG_ERROR("Export type not supported.");
        return;
    }
    exportAsPly(exportConfig["Common"], pos, color, phase);
}

void ExportUtil::exportData(const json &exportConfig, std::vector<float> &pos) {
    std::vector<float3> p(pos.size() / 3);
    std::memcpy(p.data(), pos.data(), p.size() * sizeof(float3));
    exportData(exportConfig, p);
}

// ASCII PLY in the reference's layout (ExportUtil.cpp:28-88): x y z floats,
// then (when given) colour written as the float values under `uchar` property
// names, then p1..pN phase fractions; each row ends with a space.  The file
// is <exportTargetDir>/<exportFilePrefix>.ply.
void ExportUtil::exportAsPly(const json &exportConfig, const std::vector<float3> &pos,
                             const std::vector<float3> &color, const std::vector<float> &phase) {
    const std::string dir = exportConfig["exportTargetDir"].str();
    const std::string name = exportConfig["exportFilePrefix"].str();
    const bool exportColor = !color.empty();
    const bool exportPhase = !phase.empty();
    const int phaseNum = exportPhase ? int(exportConfig["exportPhaseNum"].number()) : 0;
    if (!make_dirs(dir)) {
        VTP_LOG_ERROR("cannot create export directory " + dir);
        return;
    }
    const std::string path = dir + "/" + name + ".ply";
    std::ofstream t(path);
    if (!t) {
        VTP_LOG_ERROR("cannot write " + path);
        return;
    }
    t << "ply\n";
    t << "format ascii 1.0\n";
    t << "element vertex " << pos.size() << "\n";
    t << "property float x\n";
    t << "property float y\n";
    t << "property float z\n";
    if (exportColor) {
        t << "property uchar red\n";
        t << "property uchar green\n";
        t << "property uchar blue\n";
    }
    for (int i = 0; i < phaseNum; ++i) t << "property float p" << i + 1 << "\n";
    t << "end_header\n";
    for (size_t i = 0; i < pos.size(); ++i) {
        t << pos[i].x << " " << pos[i].y << " " << pos[i].z << " ";
        if (exportColor) t << color[i].x << " " << color[i].y << " " << color[i].z << " ";
        for (int j = 0; j < phaseNum; ++j) t << phase[i * size_t(phaseNum) + j] << " ";
        t << "\n";
    }
    VTP_LOG_INFO("Exported file: " + path);
}

bool ExportUtil::checkConfig(const json &config) {
    if (!config.contains("Common") || !config.contains("SolverRequired")) {
        VTP_LOG_ERROR("Export config missing Common / SolverRequired");
        return false;
    }
    for (const auto &k : exportConfigRequiredCommonKeys)
        if (!config["Common"].contains(k)) {
            VTP_LOG_ERROR("Export Common Config missing key: " + k);
            return false;
        }
    for (const auto &k : exportConfigRequiredSolverRequiredKeys)
        if (!config["SolverRequired"].contains(k)) {
            VTP_LOG_ERROR("Export SolverRequired Config missing key: " + k);
            return false;
        }
    return true;
}

// ---------------------------------------------------------------- components
namespace {
const char *kComponentName[6] = {"ParticleGeometryComponent", "ParticleSphereComponent", "ParticleCubeComponent",
                                 "ParticleCylinderComponent", "ParticlePlaneComponent", "ParticleBoxComponent"};
}

json ParticleGeometryComponent::getConfig() {
    json j = json::object();
    j["epmMaterial"] = int(epmMaterial);
    j["particleRadius"] = particleRadius;
    switch (m_type) {
        case Particle_Common: j["particleGeometryPath"] = ""; break;
        case Particle_Sphere:
            j["volumeCenter"] = json({0.0, 0.0, 0.0});
            j["volumeRadius"] = 1.0;
            break;
        case Particle_Cube:
            j["lb"] = json({-1.0, -1.0, -1.0});
            j["size"] = json({2.0, 2.0, 2.0});
            break;
        case Particle_Cylinder:
            j["center"] = json({0.0, 0.0, 0.0});
            j["bottomAreaRadius"] = 1.0;
            j["height"] = 2.0;
            break;
        case Particle_Plane:
            j["lb"] = json({0.0, 1.0, 0.0});
            j["size"] = json({0.0, 0.0, 0.0});
            j["layerNum"] = 2;
            break;
        case Particle_Box:
            j["lb"] = json({-1.0, -1.0, -1.0});
            j["size"] = json({2.0, 2.0, 2.0});
            j["layerNum"] = 1;
            break;
        default: break;
    }
    return j;
}

bool ParticleGeometryComponent::configIsComplete(const json &config) const {
    static const std::vector<std::string> keys[6] = {
        {"particleRadius", "particleGeometryPath"},
        {"particleRadius", "volumeCenter", "volumeRadius"},
        {"particleRadius", "lb", "size"},
        {"particleRadius", "center", "bottomAreaRadius", "height"},
        {"particleRadius", "lb", "size", "layerNum"},
        {"particleRadius", "lb", "size", "layerNum"},
    };
    const int t = int(m_type);
    if (t < 0 || t > 5) return false;
    bool ok = true;
    for (const auto &k : keys[t])
        if (!config.contains(k)) {
            VTP_LOG_ERROR(std::string(kComponentName[t]) + " config missing key: " + k);
            ok = false;
        }
    if (!config.contains("epmMaterial"))
        VTP_LOG_WARNING(std::string(kComponentName[t]) + " config missing key: epmMaterial");
    return ok;
}

void ParticleGeometryComponent::update(json &config) {
    if (!configIsComplete(config)) return;
    if (m_type == Particle_Common)
        config["loadType"] = int(m_type);
    else
        config["genType"] = int(m_type);
    if (config.contains("epmMaterial")) epmMaterial = uint8_t(int(config["epmMaterial"].number()));
    particleRadius = f(config["particleRadius"]);
    pos = m_type == Particle_Common ? ModelHandler::loadObjectElements(config)
                                    : ModelHandler::generateObjectElements(config);
    VTP_LOG_INFO(std::string(kComponentName[int(m_type)]) + " config updated.");
}

std::vector<float> ParticleGeometryComponent::getElements() {
    std::vector<float> out(pos.size() * 3);
    std::memcpy(out.data(), pos.data(), out.size() * sizeof(float));
    return out;
}

// ---------------------------------------------------------------- Object
Object::Object(long long id) : m_id(id), m_name("obj_" + std::to_string(id)) {
    VTP_LOG_INFO("Object: " + m_name + " Created.");
}

Object::~Object() { delete m_objectTypeComponent; }

void Object::transferTo(ObjectType type) {
    ObjectTypeComponent *c = nullptr;
    if (int(type) <= int(Particle_Box))
        c = new ParticleGeometryComponent(type);
    else if (type == MeshGeometry)
        c = new MeshGeometryComponent();
    if (!c) return;  // no component template of that type: unchanged
    delete m_objectTypeComponent;
    m_objectTypeComponent = c;
    m_objectComponentConfig = c->getConfig();
}

json &Object::getObjectComponentConfig() { return m_objectComponentConfig; }
ObjectTypeComponent *Object::getObjectComponent() { return m_objectTypeComponent; }
void Object::attachSpecificSolverObjectComponentConfig(json cfg) { m_solverSpecificComponentConfig = std::move(cfg); }
json &Object::getSolverObjectComponentConfig() { return m_solverSpecificComponentConfig; }
long long Object::getID() const { return m_id; }
void Object::rename(std::string newName) { m_name = std::move(newName); }
std::string Object::getName() const { return m_name; }

bool Object::update() {
    if (!m_objectTypeComponent) {
        VTP_LOG_ERROR("Object: " + m_name + " has no component (transferTo a type first).");
        return false;
    }
    m_objectTypeComponent->update(m_objectComponentConfig);
    return true;
}

void Object::reset() {
    m_objectComponentConfig = json();
    m_solverSpecificComponentConfig = json();
    delete m_objectTypeComponent;
    m_objectTypeComponent = nullptr;
}

// ---------------------------------------------------------------- ObjectManager
Object *ObjectManager::createObject(ObjectType type) {
    Object *obj = new Object(m_objectNum++);
    obj->transferTo(type);
    m_objects.push_back(obj);
    return obj;
}

void ObjectManager::clear() {
    for (Object *o : m_objects) {
        o->reset();
        delete o;
    }
    m_objects.clear();
}

std::vector<Object *> ObjectManager::createObjectsFromJson(json allConfiguration) {
    if (!allConfiguration.contains("OBJECTS") || allConfiguration["OBJECTS"].empty()) {
        VTP_LOG_ERROR("ObjectManager::createObjectsFromJson: objectsConfig is empty");
        return {};
    }
    std::vector<Object *> ret;
    for (const auto &oc : allConfiguration["OBJECTS"].items()) {
        Object *obj = createObject(ObjectType(uint8_t(int(oc["objTransferType"].number()))));
        obj->getObjectComponentConfig() = oc["ObjectComponentConfig"];
        obj->attachSpecificSolverObjectComponentConfig(oc["SolverObjectComponentConfig"]);
        obj->update();
        ret.push_back(obj);
    }
    return ret;
}

}  // namespace VT_Physics
