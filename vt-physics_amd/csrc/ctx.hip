// ctx.hip - solver context and the C ABI (include/vtp.h).
//
// Mirrors the life cycle of the reference Solver (Solver.hpp:27-62):
// create = SolverManager::createSolver + setConfig, add_particles =
// attachObject, initialize, step = tickNsteps, run, destroy.  Host arrays in
// and out are in attach order; on the device the state lives in cell-sorted
// slot order and `orig` maps slot -> attach id.
#include <algorithm>
#include <cmath>
#include <iterator>
#include <cstring>
#include <numeric>

#include "vtp_internal.h"

namespace vtp {

const char *solver_template(int solver);
const char *solver_domain(int solver);

uint32_t div_up(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

// Hash without prediction (UGNS::update on the current positions).
__global__ void k_hash(const float4 *__restrict__ x, const uint32_t *__restrict__ scene, uint32_t n, Grid g,
                       uint32_t *__restrict__ key, uint32_t *__restrict__ val, uint32_t *__restrict__ status) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 xi = x[i];
    uint32_t k = cell_key(xyz(xi), scene[i], g);
    bool finite = isfinite(xi.x) && isfinite(xi.y) && isfinite(xi.z);
    if (!finite) k = g.dead_key;
    key[i] = k;
    val[i] = i;
    if (k == g.dead_key) {
        atomicAdd(&status[kStDead], 1u);
        atomicOr(&status[kStCrash], finite ? (uint32_t)VTP_CRASH_OUT_OF_GRID : (uint32_t)VTP_CRASH_NAN);
    }
}

__global__ void k_iota(uint32_t *a, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = i;
}

// New slot i takes old slot perm[i]; static per-particle data (material,
// scene) comes from the attach-order tables through the new attach id.
// Non-fluid particles are also written to the second destination buffers,
// which lets later passes ping-pong without touching them.
__global__ void k_reorder_state(const uint32_t *__restrict__ perm, const uint32_t *__restrict__ orig_new,
                                uint32_t n, const float4 *__restrict__ xs, float4 *__restrict__ xd,
                                float4 *__restrict__ xd2, const float4 *__restrict__ vs,
                                float4 *__restrict__ vd, float4 *__restrict__ vd2,
                                const int32_t *__restrict__ mat_attach, const uint32_t *__restrict__ scene_attach,
                                const uint32_t *__restrict__ obj_attach, int32_t *__restrict__ mat,
                                uint32_t *__restrict__ scene, uint32_t *__restrict__ orig,
                                uint32_t *__restrict__ obj_of) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t o = perm[i];
    uint32_t a = orig_new[i];
    float4 x = xs[o], v = vs[o];
    int32_t m = mat_attach[a];
    xd[i] = x;
    vd[i] = v;
    if (m != VTP_MAT_FLUID) {
        if (xd2) xd2[i] = x;
        if (vd2) vd2[i] = v;
    }
    mat[i] = m;
    scene[i] = scene_attach[a];
    orig[i] = a;
    if (obj_of) obj_of[i] = obj_attach[a];
}

__global__ void k_step_begin(uint32_t *status) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        status[kStDead] = 0;
        status[kStNbrPrev] = status[kStNbrAtCap];
        status[kStNbrAtCap] = 0;
    }
}

void common_reorder(Ctx &c, const float4 *xs, float4 *xd, float4 *xd2, const float4 *vs, float4 *vd,
                    float4 *vd2) {
    hipLaunchKernelGGL(k_reorder_state, dim3(div_up(c.n, 256)), dim3(256), 0, c.stream, c.ugns.perm.p,
                       c.ugns.orig_new.p, c.n, xs, xd, xd2, vs, vd, vd2, c.mat_attach.p, c.scene_attach.p,
                       c.obj_attach.p, c.mat.p, c.scene.p, c.orig.p, c.obj_attach.p ? c.obj_of.p : nullptr);
}

void launch_step_begin(Ctx &c) { hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(64), 0, c.stream, c.status.p); }

void launch_hash(Ctx &c, const float4 *x) {
    hipLaunchKernelGGL(k_hash, dim3(div_up(c.n, 256)), dim3(256), 0, c.stream, x, c.scene.p, c.n, c.grid,
                       c.ugns.key[0].p, c.ugns.val[0].p, c.status.p);
}

// Upload the attach-order arrays, run the first neighbour search and leave
// the sorted state in X[0]/V[0] (PBFSolver.cpp:102-112).
void common_upload_and_search(Ctx &c) {
    const uint32_t n = c.n;
    c.mat_attach.alloc(n);
    c.scene_attach.alloc(n);
    c.obj_attach.alloc(n);
    c.mat.alloc(n);
    c.scene.alloc(n);
    c.orig.alloc(n);
    c.obj_of.alloc(n);
    for (int b = 0; b < 3; ++b) {
        c.X[b].alloc(n);
        c.V[b].alloc(n);
    }
    std::vector<float4> hx(n), hv(n);
    for (uint32_t i = 0; i < n; ++i) {
        hx[i] = make_float4(c.h_pos[3 * i], c.h_pos[3 * i + 1], c.h_pos[3 * i + 2], 0.f);
        hv[i] = make_float4(c.h_vel[3 * i], c.h_vel[3 * i + 1], c.h_vel[3 * i + 2], 0.f);
    }
    std::vector<uint32_t> hobj(n);
    for (uint32_t o = 0; o < c.objects.size(); ++o)
        for (uint32_t i = c.objects[o].start; i < c.objects[o].end; ++i) hobj[i] = o;
    VTP_HIP(hipMemcpyAsync(c.X[1].p, hx.data(), n * sizeof(float4), hipMemcpyHostToDevice, c.stream));
    VTP_HIP(hipMemcpyAsync(c.V[1].p, hv.data(), n * sizeof(float4), hipMemcpyHostToDevice, c.stream));
    VTP_HIP(hipMemcpyAsync(c.mat_attach.p, c.h_mat.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, c.stream));
    VTP_HIP(hipMemcpyAsync(c.scene_attach.p, c.h_scene.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice,
                           c.stream));
    VTP_HIP(hipMemcpyAsync(c.obj_attach.p, hobj.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, c.stream));
    VTP_HIP(hipMemcpyAsync(c.mat.p, c.mat_attach.p, n * sizeof(int32_t), hipMemcpyDeviceToDevice, c.stream));
    VTP_HIP(hipMemcpyAsync(c.scene.p, c.scene_attach.p, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, c.stream));
    hipLaunchKernelGGL(k_iota, dim3(div_up(n, 256)), dim3(256), 0, c.stream, c.orig.p, n);
    c.ugns.allocate(n, c.grid, c.p.max_neighbor_num);
    c.ugns.status = c.status.p;
    launch_step_begin(c);
    launch_hash(c, c.X[1].p);
    c.ugns.sort_and_rank(c.orig.p, c.stream, c.timer);
    common_reorder(c, c.X[1].p, c.X[0].p, nullptr, c.V[1].p, c.V[0].p, nullptr);
    c.ugns.build_neighbors(c.X[0].p, c.stream, c.timer);
    VTP_HIP(hipGetLastError());
    // the host staging copies must outlive the async uploads
    VTP_HIP(hipStreamSynchronize(c.stream));
}

void read_status(Ctx &c) {
    VTP_HIP(hipMemcpyAsync(c.h_status, c.status.p, kStWords * sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream));
    VTP_HIP(hipStreamSynchronize(c.stream));
}

Ctx::~Ctx() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (h_status) (void)hipHostFree(h_status);
    if (h_scal) (void)hipHostFree(h_scal);
    if (stream) (void)hipStreamDestroy(stream);
}

// Largest float d2 with sqrtf(d2) <= h (host sqrtf is IEEE correctly rounded).
static float neighbor_threshold(float h) {
    float t = h * h;
    while (std::sqrt(t) > h) t = std::nextafter(t, 0.0f);
    for (;;) {
        float nx = std::nextafter(t, INFINITY);
        if (std::sqrt(nx) <= h) t = nx; else break;
    }
    return t;
}

static void derive(Ctx &c) {
    const vtp_params &p = c.p;
    c.h = p.particle_radius * 4.0f;                        // PBFSolver.cpp:51
    c.V0 = std::pow(2.0f * p.particle_radius, 3.0f);        // PBFSolver.cpp:50 (powf)
    c.inv_dt = 1.0f / p.time_step;
    c.inv_dt2 = c.inv_dt * c.inv_dt;
    c.kern = make_cubic(c.h);
    Grid &g = c.grid;
    g.lb_x = p.sim_space_lb[0];
    g.lb_y = p.sim_space_lb[1];
    g.lb_z = p.sim_space_lb[2];
    g.cell = c.h;
    // UniformGridNeighborSearch.cpp:214-218: ceil in float
    g.gx = (int)(uint32_t)std::ceil(p.sim_space_size[0] / c.h);
    g.gy = (int)(uint32_t)std::ceil(p.sim_space_size[1] / c.h);
    g.gz = (int)(uint32_t)std::ceil(p.sim_space_size[2] / c.h);
    g.cells = (uint32_t)g.gx * (uint32_t)g.gy * (uint32_t)g.gz;
    g.nbr_d2_max = neighbor_threshold(c.h);
}

}  // namespace vtp

using namespace vtp;

static thread_local std::string g_create_error;

#define CTX_GUARD(ctx)                                                         \
    if (!(ctx)) return VTP_EINVAL;                                             \
    Ctx &c = (ctx)->c;

template <typename F>
static vtp_status guarded(Ctx &c, F &&f) {
    try {
        c.err.clear();
        return f();
    } catch (const HipError &e) {
        c.err = e.what();
        return e.code == hipErrorOutOfMemory ? VTP_ENOMEM : VTP_EHIP;
    } catch (const std::bad_alloc &) {
        c.err = "host allocation failed";
        return VTP_ENOMEM;
    } catch (const std::exception &e) {
        c.err = e.what();
        return VTP_EINVAL;
    }
}

static vtp_status ctx_fail(Ctx &c, vtp_status s, const std::string &msg) {
    c.err = msg;
    return s;
}

extern "C" vtp_status vtp_create_with_params(vtp_ctx **out, const vtp_params *params) {
    if (!out || !params) {
        g_create_error = "vtp_create: null argument";
        return VTP_EINVAL;
    }
    *out = nullptr;
    if (params->solver < 0 || params->solver > 3) {
        g_create_error = "vtp_create: unknown solver type";
        return VTP_EINVAL;
    }
    if (params->cfl_enable && params->solver != VTP_DFSPH) {
        // only the DFSPH tick adapts dt and advances the simulated time by it
        g_create_error = "vtp_create: cfl_enable exists for DFSPH only";
        return VTP_EINVAL;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        g_create_error = "vtp_create: no HIP device available";
        return VTP_ENODEV;
    }
    vtp_ctx *ctx = new (std::nothrow) vtp_ctx();
    if (!ctx) return VTP_ENOMEM;
    Ctx &c = ctx->c;
    c.p = *params;
    if (c.p.max_solver_iterations == 0) c.p.max_solver_iterations = 100;
    if (c.p.dt_max <= 0.f) c.p.dt_max = c.p.time_step;
    vtp_status st = guarded(c, [&]() {
        VTP_HIP(hipGetDevice(&c.device));
        VTP_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
        VTP_HIP(hipHostMalloc(reinterpret_cast<void **>(&c.h_status), kStWords * sizeof(uint32_t)));
        VTP_HIP(hipHostMalloc(reinterpret_cast<void **>(&c.h_scal), sizeof(SolverScalars)));
        std::memset(c.h_status, 0, kStWords * sizeof(uint32_t));
        derive(c);
        return VTP_OK;
    });
    if (st != VTP_OK) {
        g_create_error = c.err;
        delete ctx;
        return st;
    }
    *out = ctx;
    return VTP_OK;
}

extern "C" vtp_status vtp_create(vtp_ctx **out, int32_t solver, const char *json_text) {
    vtp_params p;
    char err[512] = {0};
    vtp_status s = vtp_params_from_json(solver, json_text, &p, err, sizeof err);
    if (s != VTP_OK) {
        g_create_error = err;
        if (out) *out = nullptr;
        return s;
    }
    return vtp_create_with_params(out, &p);
}

extern "C" void vtp_destroy(vtp_ctx *ctx) {
    if (!ctx) return;
    if (ctx->c.stream) (void)hipStreamSynchronize(ctx->c.stream);
    delete ctx;
}

extern "C" const char *vtp_last_error(const vtp_ctx *ctx) {
    return ctx ? ctx->c.err.c_str() : g_create_error.c_str();
}

extern "C" vtp_status vtp_add_particles(vtp_ctx *ctx, uint32_t scene, int32_t mat, const float *xyz,
                                        const float *vel, const float *vol_frac, uint32_t n) {
    CTX_GUARD(ctx);
    if (c.initialized) return ctx_fail(c, VTP_ESTATE, "vtp_add_particles: solver already initialized");
    if (!xyz && n) return ctx_fail(c, VTP_EINVAL, "vtp_add_particles: null positions");
    if (mat != VTP_MAT_FLUID && mat != VTP_MAT_BOUNDARY)
        return ctx_fail(c, VTP_EINVAL, "vtp_add_particles: unsupported material");  // PBFSolver.cpp:175-178
    if (c.objects.size() >= VTP_MAX_OBJECTS) return ctx_fail(c, VTP_EINVAL, "vtp_add_particles: too many objects");
    // neighbour gathers address a particle array with a 32-bit byte offset (16 B elements)
    if ((uint64_t)c.n + n > (1ull << 28))
        return ctx_fail(c, VTP_EINVAL, "vtp_add_particles: at most 2^28 particles per context");
    return guarded(c, [&]() {
        Object o{c.n, c.n + n, scene, mat};
        c.h_pos.insert(c.h_pos.end(), xyz, xyz + 3 * (size_t)n);
        if (vel) c.h_vel.insert(c.h_vel.end(), vel, vel + 3 * (size_t)n);
        else c.h_vel.insert(c.h_vel.end(), 3 * (size_t)n, 0.0f);
        c.h_mat.insert(c.h_mat.end(), n, mat);
        c.h_scene.insert(c.h_scene.end(), n, scene);
        const uint32_t P = (c.p.solver == VTP_IMM || c.p.solver == VTP_IMMCT) ? c.p.phase_num : 0;
        if (P) {
            if (vol_frac) c.h_frac.insert(c.h_frac.end(), vol_frac, vol_frac + (size_t)P * n);
            else c.h_frac.insert(c.h_frac.end(), (size_t)P * n, 0.0f);
        }
        c.n += n;
        c.scenes = std::max(c.scenes, scene + 1);
        c.objects.push_back(o);
        return VTP_OK;
    });
}

extern "C" vtp_status vtp_initialize(vtp_ctx *ctx) {
    CTX_GUARD(ctx);
    if (c.initialized) return ctx_fail(c, VTP_ESTATE, "vtp_initialize: already initialized");
    if (c.n == 0) return ctx_fail(c, VTP_EINVAL, "vtp_initialize: no particles attached");
    return guarded(c, [&]() {
        uint64_t keys = (uint64_t)c.scenes * c.grid.cells;
        if (keys >= 0xFFFFFFF0ull) throw std::runtime_error("grid too large: scenes*cells exceeds 32-bit keys");
        c.grid.scenes = c.scenes;
        c.grid.dead_key = (uint32_t)keys;
        c.status.alloc(kStWords);
        c.scal.alloc(1);
        VTP_HIP(hipMemsetAsync(c.status.p, 0, c.status.bytes(), c.stream));
        VTP_HIP(hipMemsetAsync(c.scal.p, 0, c.scal.bytes(), c.stream));
        switch (c.p.solver) {
            case VTP_PBF: pbf_initialize(c); break;
            case VTP_DFSPH: dfsph_initialize(c); break;
            default: imm_initialize(c); break;
        }
        read_status(c);
        c.initialized = true;
        if (c.opt_autotune) autotune(c, c.opt_autotune);   // requested before initialize
        return VTP_OK;
    });
}

static vtp_status check_after_steps(Ctx &c) {
    c.timer.collect();
    read_status(c);
    uint32_t fatal = c.h_status[kStCrash] & ~(uint32_t)VTP_CRASH_OUT_OF_GRID;
    if (fatal) {
        c.crashed = true;
        return ctx_fail(c, VTP_ECRASH, "solver crashed (flags " + std::to_string(c.h_status[kStCrash]) + ")");
    }
    return VTP_OK;
}

void vtp::solver_tick(Ctx &c) {
    c.lists_stale = false;   // every tick rebuilds the lists first
    switch (c.p.solver) {
        case VTP_PBF: pbf_step(c); break;
        case VTP_DFSPH: dfsph_step(c); break;
        default: imm_step(c); break;
    }
}

extern "C" vtp_status vtp_step(vtp_ctx *ctx, uint32_t nsteps) {
    CTX_GUARD(ctx);
    if (!c.initialized) {
        vtp_status s = vtp_initialize(ctx);
        if (s != VTP_OK) return s;
    }
    if (c.crashed) return ctx_fail(c, VTP_ECRASH, "solver is crashed");
    return guarded(c, [&]() {
        for (uint32_t s = 0; s < nsteps; ++s) {
            solver_tick(c);
            c.steps += 1;
            // a DFSPH tick with CFL advances sim_time by its own dt (dfsph_step)
            if (!(c.p.solver == VTP_DFSPH && c.p.cfl_enable)) c.sim_time += c.p.time_step;
        }
        return check_after_steps(c);
    });
}

extern "C" vtp_status vtp_run(vtp_ctx *ctx) {
    CTX_GUARD(ctx);
    if (!c.initialized) {
        vtp_status s = vtp_initialize(ctx);
        if (s != VTP_OK) return s;
    }
    while (c.sim_time < c.p.animation_time) {
        vtp_status s = vtp_step(ctx, 1);
        if (s != VTP_OK) return s;
    }
    return VTP_OK;
}

extern "C" uint32_t vtp_particle_count(const vtp_ctx *ctx) { return ctx ? ctx->c.n : 0; }

namespace vtp {
// Device slot-order array -> host attach-order array (comps values per particle).
template <typename T>
static void to_attach(Ctx &c, const T *dev, size_t comps_per_particle, std::vector<uint32_t> &orig, T *dst,
                      size_t phase_stride_elems = 0, uint32_t phases = 0) {
    (void)phase_stride_elems;
    (void)phases;
    std::vector<T> tmp((size_t)c.n * comps_per_particle);
    VTP_HIP(hipMemcpyAsync(tmp.data(), dev, tmp.size() * sizeof(T), hipMemcpyDeviceToHost, c.stream));
    VTP_HIP(hipStreamSynchronize(c.stream));
    for (uint32_t i = 0; i < c.n; ++i)
        std::memcpy(dst + (size_t)orig[i] * comps_per_particle, tmp.data() + (size_t)i * comps_per_particle,
                    comps_per_particle * sizeof(T));
}
}  // namespace vtp

static size_t field_floats(const Ctx &c, int field) {
    const size_t n = c.n, P = c.P;
    switch (field) {
        case VTP_FIELD_POS: case VTP_FIELD_VEL: case VTP_FIELD_VEL_ADV: case VTP_FIELD_ACC:
        case VTP_FIELD_SURFACE_NORMAL: case VTP_FIELD_COLOR: return 3 * n;
        case VTP_FIELD_VOL_FRAC: return P * n;
        case VTP_FIELD_VEL_PHASE: case VTP_FIELD_VEL_DRIFT: return 3 * P * n;
        case VTP_FIELD_Q: case VTP_FIELD_VEL_GRAD: case VTP_FIELD_VISE_STRESS: return 9 * n;
        default: return n;
    }
}

extern "C" vtp_status vtp_read(vtp_ctx *ctx, int32_t field, void *host_dst, size_t bytes) {
    CTX_GUARD(ctx);
    if (!c.initialized) return ctx_fail(c, VTP_ESTATE, "vtp_read: not initialized");
    if (field < 0 || field >= VTP_FIELD_COUNT_) return ctx_fail(c, VTP_EINVAL, "vtp_read: unknown field");
    if (!host_dst) return ctx_fail(c, VTP_EINVAL, "vtp_read: null destination");
    if (bytes != field_floats(c, field) * 4)
        return ctx_fail(c, VTP_EINVAL, "vtp_read: size mismatch (expected " +
                                           std::to_string(field_floats(c, field) * 4) + " bytes)");
    return guarded(c, [&]() -> vtp_status {
        const uint32_t n = c.n;
        std::vector<uint32_t> orig(n);
        VTP_HIP(hipMemcpyAsync(orig.data(), c.orig.p, n * 4, hipMemcpyDeviceToHost, c.stream));
        VTP_HIP(hipStreamSynchronize(c.stream));
        float *out = static_cast<float *>(host_dst);
        auto f4_xyz = [&](const float4 *d) {
            if (!d) throw std::runtime_error("field not available for this solver");
            std::vector<float4> t(n);
            VTP_HIP(hipMemcpyAsync(t.data(), d, n * sizeof(float4), hipMemcpyDeviceToHost, c.stream));
            VTP_HIP(hipStreamSynchronize(c.stream));
            for (uint32_t i = 0; i < n; ++i) {
                out[3 * (size_t)orig[i]] = t[i].x;
                out[3 * (size_t)orig[i] + 1] = t[i].y;
                out[3 * (size_t)orig[i] + 2] = t[i].z;
            }
        };
        auto f4_w = [&](const float4 *d) {
            std::vector<float4> t(n);
            VTP_HIP(hipMemcpyAsync(t.data(), d, n * sizeof(float4), hipMemcpyDeviceToHost, c.stream));
            VTP_HIP(hipStreamSynchronize(c.stream));
            for (uint32_t i = 0; i < n; ++i) out[orig[i]] = t[i].w;
        };
        auto f1 = [&](const float *d) {
            if (!d) throw std::runtime_error("field not available for this solver");
            to_attach(c, d, 1, orig, out);
        };
        // a table already in attach order (per-particle constants set at initialize)
        auto attach1 = [&](const float *d) {
            if (!d) throw std::runtime_error("field not available for this solver");
            VTP_HIP(hipMemcpyAsync(out, d, (size_t)n * 4, hipMemcpyDeviceToHost, c.stream));
            VTP_HIP(hipStreamSynchronize(c.stream));
        };
        // phase-major [k*stride + slot] arrays -> [attach*P + k] (comps values each)
        auto phased = [&](const void *d, size_t elem_floats, size_t comps) {
            if (!d) throw std::runtime_error("field not available for this solver");
            const size_t stride = c.ugns.stride;
            const size_t P = c.P;
            std::vector<float> t(P * stride * elem_floats);
            VTP_HIP(hipMemcpyAsync(t.data(), d, t.size() * 4, hipMemcpyDeviceToHost, c.stream));
            VTP_HIP(hipStreamSynchronize(c.stream));
            for (uint32_t i = 0; i < n; ++i)
                for (size_t k = 0; k < P; ++k)
                    for (size_t a = 0; a < comps; ++a)
                        out[((size_t)orig[i] * P + k) * comps + a] = t[(k * stride + i) * elem_floats + a];
        };
        // plus_identity: the device holds Q - I (full FP32 precision near equilibrium)
        auto mat9 = [&](const float *d, bool plus_identity = false) {
            if (!d) throw std::runtime_error("field not available for this solver");
            const size_t stride = c.ugns.stride;
            std::vector<float> t(9 * stride);
            VTP_HIP(hipMemcpyAsync(t.data(), d, t.size() * 4, hipMemcpyDeviceToHost, c.stream));
            VTP_HIP(hipStreamSynchronize(c.stream));
            for (uint32_t i = 0; i < n; ++i)
                for (size_t a = 0; a < 9; ++a)
                    out[(size_t)orig[i] * 9 + a] =
                        t[a * stride + i] + ((plus_identity && (a == 0 || a == 4 || a == 8)) ? 1.0f : 0.0f);
        };
        switch (field) {
            case VTP_FIELD_POS: f4_xyz(c.X[0].p); break;
            case VTP_FIELD_VEL: f4_xyz(c.V[0].p); break;
            case VTP_FIELD_MAT: to_attach(c, c.mat.p, 1, orig, static_cast<int32_t *>(host_dst)); break;
            case VTP_FIELD_SCENE: to_attach(c, c.scene.p, 1, orig, static_cast<uint32_t *>(host_dst)); break;
            case VTP_FIELD_MASS: f4_w(c.X[0].p); break;
            case VTP_FIELD_VOLUME: f4_w(c.V[0].p); break;
            case VTP_FIELD_DENSITY: f1(c.dens.p); break;
            case VTP_FIELD_LAMBDA: f1(c.lam.p); break;
            case VTP_FIELD_ERROR: f1(c.err_c.p); break;
            case VTP_FIELD_REST_DENSITY:
                if (c.P) f1(c.rest2[0].p); else attach1(c.rest_attach.p);
                break;
            case VTP_FIELD_COMPRESSION_RATIO: f1(c.cr.p); break;
            case VTP_FIELD_DF_ALPHA: f1(c.alpha.p); break;
            case VTP_FIELD_DELTA_CR: f1(c.delta.p); break;
            case VTP_FIELD_KAPPA_DIV: f1(c.kappa_div.p); break;
            case VTP_FIELD_KAPPA_INCOMP: f1(c.kappa.p); break;
            case VTP_FIELD_VEL_ADV: f4_xyz(c.vadv.p); break;
            case VTP_FIELD_ACC: f4_xyz(c.acc.p); break;
            case VTP_FIELD_SURFACE_NORMAL: f4_xyz(c.normal.p); break;
            case VTP_FIELD_VIS:
                if (c.p.solver == VTP_IMMCT) f1(c.vis_ct.p); else if (c.P) f1(c.vis.p); else attach1(c.vis_attach.p);
                break;
            case VTP_FIELD_VOL_FRAC: phased(c.frac[0].p, 1, 1); break;
            case VTP_FIELD_VEL_PHASE: phased(c.uphase[0].p, 4, 3); break;
            case VTP_FIELD_VEL_DRIFT: phased(c.drift.p, 4, 3); break;
            case VTP_FIELD_COLOR: f4_xyz(c.P ? c.color4[0].p : c.color.p); break;
            case VTP_FIELD_CD: f1(c.P ? c.cd.p : nullptr); break;
            case VTP_FIELD_Q: case VTP_FIELD_VEL_GRAD: case VTP_FIELD_VISE_STRESS:
                if (c.p.solver != VTP_IMMCT) throw std::runtime_error("field only exists for IMM-CT");
                mat9(field == VTP_FIELD_Q ? c.Q[0].p : field == VTP_FIELD_VEL_GRAD ? c.Lgrad.p : c.stress.p,
                     field == VTP_FIELD_Q);
                break;
            default: throw std::runtime_error("field not available");
        }
        return VTP_OK;
    });
}

extern "C" vtp_status vtp_write(vtp_ctx *ctx, int32_t field, const void *host_src, size_t bytes) {
    CTX_GUARD(ctx);
    if (!c.initialized) return ctx_fail(c, VTP_ESTATE, "vtp_write: not initialized");
    if (field != VTP_FIELD_POS && field != VTP_FIELD_VEL)
        return ctx_fail(c, VTP_EINVAL, "vtp_write: only POS and VEL are writable");
    if (!host_src || bytes != (size_t)c.n * 12) return ctx_fail(c, VTP_EINVAL, "vtp_write: size mismatch");
    return guarded(c, [&]() -> vtp_status {
        const uint32_t n = c.n;
        std::vector<uint32_t> orig(n);
        std::vector<float4> t(n);
        float4 *dst = field == VTP_FIELD_POS ? c.X[0].p : c.V[0].p;
        VTP_HIP(hipMemcpyAsync(orig.data(), c.orig.p, n * 4, hipMemcpyDeviceToHost, c.stream));
        VTP_HIP(hipMemcpyAsync(t.data(), dst, n * sizeof(float4), hipMemcpyDeviceToHost, c.stream));
        VTP_HIP(hipStreamSynchronize(c.stream));
        const float *src = static_cast<const float *>(host_src);
        for (uint32_t i = 0; i < n; ++i) {
            size_t a = orig[i];
            t[i].x = src[3 * a];
            t[i].y = src[3 * a + 1];
            t[i].z = src[3 * a + 2];
        }
        VTP_HIP(hipMemcpyAsync(dst, t.data(), n * sizeof(float4), hipMemcpyHostToDevice, c.stream));
        if (field == VTP_FIELD_VEL && c.vadv.p)
            VTP_HIP(hipMemcpyAsync(c.vadv.p, t.data(), n * sizeof(float4), hipMemcpyHostToDevice, c.stream));
        VTP_HIP(hipStreamSynchronize(c.stream));
        return VTP_OK;
    });
}

namespace vtp {
// xyz of a slot-order float4 field to attach-order float[n][3], and back
// (keeping w; `dst2`, if given, receives the same xyz: v* follows v).
__global__ void k_xyz_to_attach(const float4 *__restrict__ src, const uint32_t *__restrict__ orig, uint32_t n,
                                float *__restrict__ dst) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 v = src[i];
    const size_t o = 3 * (size_t)orig[i];
    dst[o] = v.x;
    dst[o + 1] = v.y;
    dst[o + 2] = v.z;
}
__global__ void k_xyz_from_attach(const float *__restrict__ src, const uint32_t *__restrict__ orig, uint32_t n,
                                  float4 *__restrict__ dst, float4 *__restrict__ dst2) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t o = 3 * (size_t)orig[i];
    float4 v = dst[i];
    v.x = src[o];
    v.y = src[o + 1];
    v.z = src[o + 2];
    dst[i] = v;
    if (dst2) dst2[i] = v;
}
}  // namespace vtp

extern "C" vtp_status vtp_read_device(vtp_ctx *ctx, int32_t field, float *d_dst, size_t bytes) {
    CTX_GUARD(ctx);
    if (!c.initialized) return ctx_fail(c, VTP_ESTATE, "vtp_read_device: not initialized");
    if (field != VTP_FIELD_POS && field != VTP_FIELD_VEL)
        return ctx_fail(c, VTP_EINVAL, "vtp_read_device: only POS and VEL");
    if (!d_dst || bytes != (size_t)c.n * 12) return ctx_fail(c, VTP_EINVAL, "vtp_read_device: size mismatch");
    return guarded(c, [&]() -> vtp_status {
        hipLaunchKernelGGL(k_xyz_to_attach, dim3(div_up(c.n, 256)), dim3(256), 0, c.stream,
                           field == VTP_FIELD_POS ? c.X[0].p : c.V[0].p, c.orig.p, c.n, d_dst);
        VTP_HIP(hipGetLastError());
        VTP_HIP(hipStreamSynchronize(c.stream));
        return VTP_OK;
    });
}

extern "C" vtp_status vtp_write_device(vtp_ctx *ctx, int32_t field, const float *d_src, size_t bytes) {
    CTX_GUARD(ctx);
    if (!c.initialized) return ctx_fail(c, VTP_ESTATE, "vtp_write_device: not initialized");
    if (field != VTP_FIELD_POS && field != VTP_FIELD_VEL)
        return ctx_fail(c, VTP_EINVAL, "vtp_write_device: only POS and VEL are writable");
    if (!d_src || bytes != (size_t)c.n * 12) return ctx_fail(c, VTP_EINVAL, "vtp_write_device: size mismatch");
    return guarded(c, [&]() -> vtp_status {
        float4 *dst = field == VTP_FIELD_POS ? c.X[0].p : c.V[0].p;
        float4 *dst2 = field == VTP_FIELD_VEL ? c.vadv.p : nullptr;
        hipLaunchKernelGGL(k_xyz_from_attach, dim3(div_up(c.n, 256)), dim3(256), 0, c.stream, d_src, c.orig.p, c.n,
                           dst, dst2);
        VTP_HIP(hipGetLastError());
        VTP_HIP(hipStreamSynchronize(c.stream));
        return VTP_OK;
    });
}

extern "C" vtp_status vtp_stats(vtp_ctx *ctx, vtp_step_stats *out) {
    CTX_GUARD(ctx);
    if (!out) return ctx_fail(c, VTP_EINVAL, "vtp_stats: null output");
    return guarded(c, [&]() -> vtp_status {
        std::memset(out, 0, sizeof *out);
        out->steps = c.steps;
        out->sim_time = c.sim_time;
        out->dt = c.p.time_step;
        if (c.initialized) {
            read_status(c);
            out->dead_particles = c.h_status[kStDead];
            out->crash_flags = c.h_status[kStCrash];
            out->div_iterations = c.h_status[kStDivIter];
            out->incomp_iterations = c.h_status[kStIncIter];
            out->phase_iterations = c.h_status[kStPhaseIter];
            if (c.p.solver != VTP_PBF) {
                VTP_HIP(hipMemcpyAsync(c.h_scal, c.scal.p, sizeof(SolverScalars), hipMemcpyDeviceToHost, c.stream));
                VTP_HIP(hipStreamSynchronize(c.stream));
                out->div_residual = c.h_scal->div_residual;
                out->incomp_residual = c.h_scal->inc_residual;
                out->dt = c.h_scal->dt;
            }
        }
        out->total_div_iterations = c.total_div;
        out->total_incomp_iterations = c.total_inc;
        return VTP_OK;
    });
}

extern "C" vtp_status vtp_debug_neighbors(vtp_ctx *ctx, uint32_t *neighbors, uint32_t *neighbor_num,
                                          uint32_t *sorted_ids, uint32_t *cell_start, uint32_t *cell_end) {
    CTX_GUARD(ctx);
    if (!c.initialized) return ctx_fail(c, VTP_ESTATE, "vtp_debug_neighbors: not initialized");
    if (c.lists_stale)
        return ctx_fail(c, VTP_ESTATE, "vtp_debug_neighbors: lists belong to a replaced state; tick first");
    if (!c.ugns.static_rows)
        return ctx_fail(c, VTP_ESTATE,
                        "vtp_debug_neighbors: the last search built fluid lists only (PBF and DFSPH ticks); set option "
                        "static_lists=1 and tick for the reference's complete lists");
    return guarded(c, [&]() -> vtp_status {
        const uint32_t n = c.n, M = c.p.max_neighbor_num;
        if (neighbors || neighbor_num) {
            DevBuf<uint32_t> lists, num;
            lists.alloc((size_t)n * M);
            num.alloc(n);
            c.ugns.export_reference(c.orig.p, lists.p, num.p, c.stream);
            VTP_HIP(hipGetLastError());
            if (neighbors)
                VTP_HIP(hipMemcpyAsync(neighbors, lists.p, (size_t)n * M * 4, hipMemcpyDeviceToHost, c.stream));
            if (neighbor_num) VTP_HIP(hipMemcpyAsync(neighbor_num, num.p, (size_t)n * 4, hipMemcpyDeviceToHost, c.stream));
            VTP_HIP(hipStreamSynchronize(c.stream));
        }
        if (sorted_ids)
            VTP_HIP(hipMemcpyAsync(sorted_ids, c.orig.p, (size_t)n * 4, hipMemcpyDeviceToHost, c.stream));
        const size_t cells = (size_t)c.grid.cells * c.scenes;
        if (cell_start)
            VTP_HIP(hipMemcpyAsync(cell_start, c.ugns.start.p, cells * 4, hipMemcpyDeviceToHost, c.stream));
        if (cell_end) VTP_HIP(hipMemcpyAsync(cell_end, c.ugns.end.p, cells * 4, hipMemcpyDeviceToHost, c.stream));
        VTP_HIP(hipStreamSynchronize(c.stream));
        return VTP_OK;
    });
}

extern "C" vtp_status vtp_grid_info(const vtp_ctx *ctx, uint32_t grid[3], uint32_t *cells) {
    if (!ctx) return VTP_EINVAL;
    const Ctx &c = ctx->c;
    if (grid) {
        grid[0] = (uint32_t)c.grid.gx;
        grid[1] = (uint32_t)c.grid.gy;
        grid[2] = (uint32_t)c.grid.gz;
    }
    if (cells) *cells = c.grid.cells;
    return VTP_OK;
}

extern "C" vtp_status vtp_set_option(vtp_ctx *ctx, const char *key, double value) {
    CTX_GUARD(ctx);
    if (!key) return ctx_fail(c, VTP_EINVAL, "vtp_set_option: null key");
    std::string k(key);
    return guarded(c, [&]() -> vtp_status {
        if (k == "graph") {
            c.opt_graph = value != 0.0;
            pbf_capture_reset(c);
        } else if (k == "autotune") {
            if (value != 0.0 && value != 1.0 && value != 2.0)
                throw std::runtime_error("autotune must be 0, 1 (workgroup sizes) or 2 (also lanes per particle)");
            c.opt_autotune = (int)value;
            if (c.opt_autotune && c.initialized) autotune(c, c.opt_autotune);
            pbf_capture_reset(c);
        } else if (k == "block") {
            uint32_t b = (uint32_t)value;
            if (b != 0 && (b % 64 != 0 || b > 256)) throw std::runtime_error("block must be a multiple of 64, <= 256");
            c.opt_block = b ? b : 256;
            c.block_of.clear();
            pbf_capture_reset(c);
        } else if (k == "stage_timing") {
            c.timer.enabled = value != 0.0;
        } else if (k == "solver_chunk") {
            c.opt_chunk = (uint32_t)value;
        } else if (k == "lanes") {
            if (value != 1.0 && value != 2.0 && value != 4.0) throw std::runtime_error("lanes must be 1, 2 or 4");
            c.opt_lanes = (uint32_t)value;
            pbf_capture_reset(c);
        } else if (k == "static_lists") {
            c.opt_static_lists = value != 0.0;
            pbf_capture_reset(c);
        } else if (k == "full_flux_retry") {
            c.opt_full_flux_retry = value != 0.0;
        } else if (k == "build_lambda") {
            c.opt_build_lambda = value != 0.0;
            pbf_capture_reset(c);
        } else if (k == "radix_sort") {
            c.ugns.radix = value != 0.0;
            pbf_capture_reset(c);
        } else if (k == "neighbor_walk") {
            if (value != 0.0 && value != 1.0 && value != 2.0)
                throw std::runtime_error("neighbor_walk must be 0 (adaptive), 1 (row) or 2 (reference)");
            c.ugns.walk = (uint32_t)value;
            pbf_capture_reset(c);
        } else {
            throw std::runtime_error("unknown option '" + k + "'");
        }
        return VTP_OK;
    });
}

extern "C" vtp_status vtp_get_option(vtp_ctx *ctx, const char *key, double *value) {
    CTX_GUARD(ctx);
    if (!key || !value) return ctx_fail(c, VTP_EINVAL, "vtp_get_option: null argument");
    const std::string k(key);
    if (k == "graph") *value = c.opt_graph;
    else if (k == "autotune") *value = c.opt_autotune;
    else if (k == "block") *value = c.opt_block;
    else if (k == "lanes") *value = c.opt_lanes;
    else if (k == "static_lists") *value = c.opt_static_lists;
    else if (k == "build_lambda") *value = c.opt_build_lambda;
    else if (k == "radix_sort") *value = c.ugns.radix;
    else if (k == "neighbor_walk") *value = c.ugns.walk;
    else if (k.rfind("block.", 0) == 0) {
        static const char *const passes[] = {"pbf_lambda", "pbf_dx", "pbf_xsph", "df_precompute", "df_delta_kappa",
                                              "df_update_vadv", "df_vis_normal", "imm_precompute", "imm_vis_normal",
                                              "imm_tension_phase_vel", "imm_fluxes", "immct_conformation",
                                              "immct_stress_acc", "imm_pass"};
        const std::string pass = k.substr(6);
        if (std::find(std::begin(passes), std::end(passes), pass) == std::end(passes))
            return ctx_fail(c, VTP_EINVAL, "vtp_get_option: unknown neighbour pass '" + pass + "'");
        // the IMM passes fall back to the "imm_pass" size, the others to "block"
        auto it = c.block_of.find(pass);
        *value = it != c.block_of.end() ? it->second
                 : pass.rfind("imm", 0) == 0 ? c.block("imm_pass") : c.opt_block;
    }
    else return ctx_fail(c, VTP_EINVAL, "vtp_get_option: unknown option '" + k + "'");
    return VTP_OK;
}

extern "C" int32_t vtp_stage_times(vtp_ctx *ctx, char *names, size_t names_len, double *ms, uint64_t *launches,
                                   int32_t max_stages) {
    if (!ctx) return -1;
    Ctx &c = ctx->c;
    try {
        c.timer.collect();
    } catch (const std::exception &e) {
        c.err = e.what();
        return -1;
    }
    std::string all;
    int32_t k = 0;
    for (const auto &name : c.timer.order) {
        if (k < max_stages) {
            if (ms) ms[k] = c.timer.ms[name];
            if (launches) launches[k] = c.timer.launches[name];
        }
        all += name;
        all += '\n';
        ++k;
    }
    if (names && names_len) {
        std::strncpy(names, all.c_str(), names_len - 1);
        names[names_len - 1] = '\0';
    }
    return k;
}

extern "C" vtp_status vtp_reset_stage_times(vtp_ctx *ctx) {
    CTX_GUARD(ctx);
    return guarded(c, [&]() { c.timer.reset(); return VTP_OK; });
}

extern "C" const char *vtp_version(void) { return "vtp-mi355x 0.1 (gfx950)"; }
