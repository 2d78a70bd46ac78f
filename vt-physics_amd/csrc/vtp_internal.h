// vtp_internal.h - host-side types shared by the HIP translation units:
// device buffers, stage timing, the uniform-grid search object and the
// solver context behind the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/vtp.h"
#include "vtp_device.h"

namespace vtp {

struct HipError : std::runtime_error {
    hipError_t code;
    HipError(hipError_t c, const std::string &what) : std::runtime_error(what), code(c) {}
};

#define VTP_HIP(call)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            throw ::vtp::HipError(e_, std::string(#call) + ": " + hipGetErrorString(e_));   \
    } while (0)

// Owning device allocation.
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        release();
        if (count == 0) count = 1;
        hipError_t e = hipMalloc(reinterpret_cast<void **>(&p), count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            throw HipError(e, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed");
        }
        n = count;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    size_t bytes() const { return n * sizeof(T); }
};

// hipEvent pairs around stages, resolved lazily.  Disabled unless the
// "stage_timing" option is set; never active while a graph is captured.
struct StageTimer {
    bool enabled = false;
    struct Pending { std::string name; hipEvent_t a, b; };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    std::map<std::string, double> ms;
    std::map<std::string, uint64_t> launches;
    std::vector<std::string> order;

    hipEvent_t get() {
        if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
        hipEvent_t e;
        VTP_HIP(hipEventCreate(&e));
        return e;
    }
    void begin(hipStream_t s, const char *name) {
        if (!enabled) return;
        Pending p{name, get(), nullptr};
        VTP_HIP(hipEventRecord(p.a, s));
        pending.push_back(p);
    }
    void end(hipStream_t s) {
        if (!enabled || pending.empty()) return;
        pending.back().b = get();
        VTP_HIP(hipEventRecord(pending.back().b, s));
    }
    void collect() {
        for (auto &p : pending) {
            if (!p.b) continue;
            VTP_HIP(hipEventSynchronize(p.b));
            float t = 0;
            VTP_HIP(hipEventElapsedTime(&t, p.a, p.b));
            if (!ms.count(p.name)) order.push_back(p.name);
            ms[p.name] += t;
            launches[p.name] += 1;
            pool.push_back(p.a);
            pool.push_back(p.b);
        }
        pending.clear();
    }
    // Solver loops enqueue iterations past their exit (they return at once);
    // a loop marks the pending list before it and drops what followed the
    // last iteration that did work, so that only working launches are timed.
    size_t mark() const { return pending.size(); }
    void drop_from(size_t m) {
        for (size_t k = m; k < pending.size(); ++k) {
            pool.push_back(pending[k].a);
            if (pending[k].b) pool.push_back(pending[k].b);
        }
        if (m < pending.size()) pending.resize(m);
    }
    void reset() { collect(); ms.clear(); launches.clear(); order.clear(); }
    ~StageTimer() {
        for (auto &p : pending) { (void)hipEventDestroy(p.a); if (p.b) (void)hipEventDestroy(p.b); }
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

// Uniform-grid neighbour search state (UniformGridNeighborSearch.hpp:28-63).
struct PbfLambdaOut;

struct Ugns {
    Grid g{};
    uint32_t n = 0, M = 0, cap = 0, stride = 0;
    uint32_t key_bits = 0, passes = 0, tiles = 0;
    int sorted = 0;                 // which key/val buffer holds the sorted result
    DevBuf<uint32_t> key[2], val[2], spine, dtot;      // LSD radix sort (radix = true)
    DevBuf<uint32_t> count, rank, bsum, tile_dead, meta; // counting sort by cell (default)
    bool radix = false;
    DevBuf<uint32_t> start, end;    // [dead_key + 1], inclusive end, kEmpty = empty
    DevBuf<uint32_t> so, perm, orig_new;
    DevBuf<uint32_t> nbr, cnt;      // ELL lists nbr[t*stride + slot], count word per slot
    uint32_t *status = nullptr;     // the context's status words (cap counts)
    uint32_t walk = 0;              // list walk: 0 adaptive, 1 row walk, 2 reference walk
    bool static_rows = true;        // the last build holds the lists of static particles too

    void allocate(uint32_t n, const Grid &g, uint32_t max_neighbors);
    void sort_and_rank(const uint32_t *orig_old, hipStream_t s, StageTimer &tm);
    // fluid_only = per-slot materials: the lists of non-fluid particles are
    // left empty (count 0); nullptr builds every list, as the reference does.
    // lam: also compute the first density + lambda pass of a PBF tick from
    // the neighbours as they are listed, and store it there.
    void build_neighbors(const float4 *x, hipStream_t s, StageTimer &tm, const int32_t *fluid_only = nullptr,
                         const PbfLambdaOut *lam = nullptr);
    void export_reference(const uint32_t *orig, uint32_t *d_lists, uint32_t *d_num, hipStream_t s);
    const uint32_t *sorted_key() const { return key[sorted].p; }
};

void launch_fill_u32(uint32_t *a, uint32_t n, uint32_t v, hipStream_t s);

// Indices into the device status word array.
enum StatusWord : uint32_t {
    kStDead = 0,         // particles off the grid in the last search
    kStCrash = 1,        // VTP_CRASH_* bits
    kStDivIter = 2,      // iterations of the last divergence solve
    kStIncIter = 3,      // iterations of the last density solve
    kStPhaseIter = 4,    // iterations of the last negative-fraction loop
    kStDone = 5,         // solver-loop convergence flag
    kStIter = 6,         // running iteration counter of the active loop
    kStNegCount = 7,     // newly flagged negative fractions (IMM)
    kStNbrAtCap = 8,     // particles whose list reached the cap in the running build
    kStNbrPrev = 9,      // the same count of the previous build (steers the walk)
    kStCapSeen = 10,     // particles at the cap summed over builds since the host cleared it
                         // (the adjoint's rollout checks that no list was cut)
    kStWords = 16
};

// Device-side scalars of the iterative solvers (written by kernels).
struct SolverScalars {
    double obj_sum[VTP_MAX_OBJECTS];   // per-object residual sums of the current iteration
    double div_residual;
    double inc_residual;
    float dt;                          // current step size (CFL may change it)
    float vmax;                        // max fluid speed (CFL)
    uint32_t vmax_bits;
    uint32_t pad;
};

struct Object {
    uint32_t start, end, scene;
    int32_t mat;
};

// PBF constants passed by value to the kernels.
struct PbfK {
    Cubic k;
    float dt, inv_dt;
    float rho0, inv_rho0;
    float xsph_k;
    float gx, gy, gz;
    int mass_weighted;   // extension: m_j in the constraint gradient
};

// Where the first density + lambda pass of a PBF tick stores its results
// when the neighbour build computes it (Ugns::build_neighbors).
struct PbfLambdaOut {
    PbfK q;
    const int32_t *mat;
    float *lam, *dens, *errc;
    float4 *sg, *xl;   // sg may be null (untaped tick)
};

// One neighbour's terms of the density and lambda sums (PBFCudaImpl.cu:70-122).
// The lambda pass and the build call it per neighbour in list order, so both
// produce the same bits.  The multiply-adds are spelled out as fmaf: in the
// build the products dx*dx.. also feed the uncontracted cut-off test, and
// left to the compiler the two contexts could contract r2 differently.
template <bool MW>
__device__ __forceinline__ void pbf_lambda_pair(const float4 &xi, const float4 &xj, const Cubic &k, float &rho,
                                                float &gsx, float &gsy, float &gsz, float &s2) {
    float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
    float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    float W, f;
    cubic_w_grad(r2, k, W, f);
    rho = fmaf(xj.w, W, rho);
    if (MW) f *= xj.w;
    gsx = fmaf(dx, f, gsx);
    gsy = fmaf(dy, f, gsy);
    gsz = fmaf(dz, f, gsz);
    s2 = fmaf(f * f, r2, s2);   // |f r_vec|^2
}

//   rho_i = max(rho, rho0),  C_i = rho_i/rho0 - 1,
//   S_i = (sum_j |gradW_ij|^2 + |sum_j gradW_ij|^2) / rho0^2,  lambda_i = -C_i / (S_i + 1e-6).
__device__ __forceinline__ void pbf_lambda_store(uint32_t i, const float4 &xi, float rho, float gsx, float gsy,
                                                 float gsz, float s2, const PbfK &q, float *lam, float *dens,
                                                 float *errc, float4 *sg, float4 *xl) {
    rho = fmaxf(rho, q.rho0);
    // rho/rho0 - 1 with a rounded division, as the reference (PBFCudaImpl.cu:92):
    // a clamped particle gets exactly C = 0 (a fused rho*(1/rho0) - 1 would
    // leave 1000*0.001f - 1 = 4.7e-8, and the adjoint tests C > 0 for the clamp).
    float C = __fdiv_rn(rho, q.rho0) - 1.0f;
    float ir2 = q.inv_rho0 * q.inv_rho0;
    float S = fmaf(gsz, gsz, fmaf(gsy, gsy, fmaf(gsx, gsx, s2))) * ir2;
    const float li = -C / (S + 1e-6f);
    lam[i] = li;
    xl[i] = make_float4(xi.x, xi.y, xi.z, li);   // the delta-x pass gathers (x_j, lambda_j) in one load
    dens[i] = rho;
    errc[i] = C;
    if (sg) sg[i] = make_float4(gsx, gsy, gsz, S);   // taped for the adjoint
}

// A static particle's outputs: lambda = 0 (the reference never writes it).
__device__ __forceinline__ void pbf_lambda_store_static(uint32_t i, const float4 &xi, float *lam, float *dens,
                                                        float *errc, float4 *sg, float4 *xl) {
    lam[i] = 0.0f;
    dens[i] = 0.0f;
    errc[i] = 0.0f;
    xl[i] = make_float4(xi.x, xi.y, xi.z, 0.0f);
    if (sg) sg[i] = make_float4(0, 0, 0, 0);
}

// Device buffers one taped PBF tick writes (owned by the adjoint state).
struct PbfTapeStep {
    uint32_t *perm, *orig, *cnt, *nbr;
    float4 *x[17];       // positions entering iteration m (m < K) and after the last one (m = K)
    float *lam[16], *C[16];
    float4 *sg[16];      // (sum_j gradW_ij, S_i) per iteration
    float4 *v_last;      // velocities entering XSPH
};

// Particle count below which neighbour passes take 4 lanes per particle:
// with one lane, such a scene leaves most CUs idle behind each particle's
// chain of dependent gathers.
constexpr uint32_t kSmallScene = 1u << 18;

struct Ctx;
PbfK pbf_constants(const Ctx &c);
void pbf_step_taped(Ctx &c, PbfTapeStep &tape);
void pbf_initialize(Ctx &c);
void pbf_step(Ctx &c);
void pbf_capture_reset(Ctx &c);
void dfsph_initialize(Ctx &c);
void dfsph_step(Ctx &c);
void imm_initialize(Ctx &c);
void imm_step(Ctx &c);
void autotune(Ctx &c, int mode);
// state images (the vtp_save_state / vtp_load_state layout), host memory
size_t state_image_bytes(Ctx &c);
vtp_status save_state_image(Ctx &c, void *host_dst, size_t bytes);
vtp_status load_state_image(Ctx &c, const void *host_src, size_t bytes);
void solver_tick(Ctx &c);

// Solver context: everything one vtp_ctx owns.
struct Ctx {
    vtp_params p{};
    std::string err;
    hipStream_t stream = nullptr;
    int device = 0;
    bool initialized = false;
    bool crashed = false;
    uint64_t steps = 0;
    float sim_time = 0;
    uint64_t total_div = 0, total_inc = 0;

    // host staging in attach order
    std::vector<float> h_pos, h_vel, h_frac;
    std::vector<int32_t> h_mat;
    std::vector<uint32_t> h_scene;
    std::vector<Object> objects;
    uint32_t n = 0, scenes = 1;

    // derived constants
    float h = 0, V0 = 0, inv_dt = 0, inv_dt2 = 0;
    Cubic kern{};
    Grid grid{};

    // options
    bool opt_graph = false;
    int opt_autotune = 0;     // 0 off, 1 workgroup sizes, 2 also PBF lanes per particle
    uint32_t opt_block = 256;
    uint32_t opt_chunk = 0;   // 0 = adaptive
    uint32_t opt_lanes = 1;   // lanes per particle in the PBF passes and the pressure solves (1, 2 or 4)
    bool opt_static_lists = false;   // PBF ticks also build the (unread) lists of static particles
    bool opt_full_flux_retry = false;  // IMM: every negative-fraction retry recomputes all fluxes
    // PBF: the list build also computes the first lambda pass of a tick.  Off
    // by default: with 16-bit lists a lambda launch is cheaper than what the
    // sums cost the build (80 registers, 6 waves per SIMD against 8).
    bool opt_build_lambda = false;
    StageTimer timer;

    // attach-order tables (static)
    DevBuf<int32_t> mat_attach;
    DevBuf<uint32_t> scene_attach;
    // slot-order state
    DevBuf<int32_t> mat;        // material per slot (refreshed from mat_attach)
    DevBuf<uint32_t> scene;     // scene per slot
    DevBuf<uint32_t> orig;      // attach id per slot
    DevBuf<float4> X[3];        // positions (w = mass), X[0] = state
    DevBuf<float4> V[3];        // velocities (w = volume), V[0] = state
    Ugns ugns;
    DevBuf<uint32_t> status;    // StatusWord
    DevBuf<SolverScalars> scal;
    uint32_t *h_status = nullptr;   // pinned mirror
    SolverScalars *h_scal = nullptr;

    // PBF scratch (slot order)
    DevBuf<float> lam, dens, err_c;
    DevBuf<float4> xl;          // (x, lambda) of the current constraint iteration
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    uint64_t capture_epoch = 0;   // bumped whenever captured kernel arguments go stale
    // PBF adjoint state (checkpoints and tapes), owned here
    std::shared_ptr<void> adj_state;
    bool lists_stale = false;   // neighbour lists do not match X[0] (after a rollout backward)

    // DFSPH / IMM fields (slot order)
    DevBuf<float> rest_density, vis, cr, alpha, delta;
    DevBuf<float> kappa_div, kappa;          // kappa of the divergence solve / of the density solve
    DevBuf<float> rest_attach, vis_attach;   // static per-particle properties, attach order
    DevBuf<uint32_t> obj_cnt;                // particles per attach object
    DevBuf<int32_t> obj_fluid;               // 1 if the object is fluid
    DevBuf<double> obj_w;                    // residual weight per object
    DevBuf<double> partial;                  // residual partial per workgroup
    uint32_t iter_hint_div = 1, iter_hint_inc = 1, iter_hint_phase = 2;
    DevBuf<float4> vadv, acc, normal;
    DevBuf<float4> xk;   // (x, V kappa) of the current solver iteration, gathered by the v* pass
    DevBuf<float4> g1;   // sum_j V_j gradW_ij of the tick (positions are fixed during the solves)
    // IMM (phase-major: [k*stride + slot])
    uint32_t P = 0;
    DevBuf<float> frac[2], frac_in, frac_out;
    DevBuf<float4> uphase[2], aphase, drift;
    DevBuf<uint32_t> flag;
    DevBuf<float4> sconst;      // DFSPH/IMM solves: (cr, alpha, V, object id bits) per slot, alpha = 0 static
    DevBuf<float4> xveff;       // IMM: (x, V) of an unflagged fluid slot, (x, 0) otherwise (phase fluxes)
    DevBuf<float> rs;           // IMM: rest density of a fluid slot, its negation for a static one
    DevBuf<float4> xmr;         // IMM: (x, m / rs) per slot, > 0 exactly for fluid (per tick)
    DevBuf<float4> aux;         // IMM: (V, polymer fraction, material bits, 0) per slot (per tick)
    DevBuf<float> phase_rho, phase_vis;
    DevBuf<float4> phase_color;
    DevBuf<float4> color;
    DevBuf<uint32_t> obj_of;   // object index per slot
    DevBuf<uint32_t> obj_attach;
    DevBuf<float> cd, blocking, fpoly;
    DevBuf<float4> color4[2];    // IMM colour state (ping-pong)
    DevBuf<float> rest2[2];      // IMM rest-density state (ping-pong)
    DevBuf<float> Q[2];          // 9 per particle, [c*stride + slot]
    DevBuf<float> Lgrad, stress; // 9 per particle
    DevBuf<float4> srec;         // IMM-CT: S / rest^2 of a slot as three float4 (fluid flag in [2].y)
    DevBuf<float> vis_ct, rheo_beta, rheo_kappa;

    // per-iteration launch shapes picked by the autotuner
    std::map<std::string, uint32_t> block_of;
    uint32_t block(const char *k) const {
        auto it = block_of.find(k);
        return it == block_of.end() ? opt_block : it->second;
    }

    ~Ctx();
};

uint32_t div_up(uint32_t a, uint32_t b);
void read_status(Ctx &c);

}  // namespace vtp

// The opaque handle of the C ABI.
struct vtp_ctx {
    vtp::Ctx c;
};
