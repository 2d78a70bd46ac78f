// pbf.hip - position-based fluids step (PBFCudaImpl.cu:34-281, PBFSolver.cpp:326-376).
//
// One tick = predict+hash -> sort -> reorder -> neighbour lists ->
// K x (density+lambda pass, delta-x pass) -> XSPH pass.  Compared with the
// reference's 8 kernels:
//   * density/error and lambda share one neighbour pass (both only read
//     positions and masses); results accumulate in registers, not in global
//     memory (the reference does `d_data->x[p_i] += ...` per neighbour);
//   * delta-x and apply_dx are one pass writing into a second position
//     buffer (Jacobi semantics preserved: every delta-x reads the positions
//     of the previous iteration), XSPH and its apply likewise;
//   * the predict (apply_ext_force) is fused with the cell hash;
//   * the first iteration's density+lambda pass is computed by the neighbour
//     build (one lane per particle), which holds every listed neighbour's
//     position and mass as it writes the list.
// Boundary particles never move: their positions sit in both ping-pong
// buffers, so the passes skip them entirely.
#include "vtp_internal.h"

namespace vtp {

PbfK pbf_constants(const Ctx &c) {
    PbfK q;
    q.k = c.kern;
    q.dt = c.p.time_step;
    q.inv_dt = c.inv_dt;
    q.rho0 = c.p.f_rest_density;
    q.inv_rho0 = 1.0f / c.p.f_rest_density;
    q.xsph_k = c.p.xsph_k;
    q.gx = c.p.gravity[0];
    q.gy = c.p.gravity[1];
    q.gz = c.p.gravity[2];
    q.mass_weighted = c.p.mass_weighted;
    return q;
}

// apply_ext_force_cuda (PBFCudaImpl.cu:146-157) + calcParticleHashValue.
__global__ void k_pbf_predict_hash(float4 *__restrict__ x, float4 *__restrict__ v,
                                   const int32_t *__restrict__ mat, const uint32_t *__restrict__ scene,
                                   uint32_t n, PbfK q, Grid g, uint32_t *__restrict__ key,
                                   uint32_t *__restrict__ val, uint32_t *__restrict__ status) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 xi = x[i];
    if (mat[i] == VTP_MAT_FLUID) {
        float4 vi = v[i];
        vi.x += q.gx * q.dt;
        vi.y += q.gy * q.dt;
        vi.z += q.gz * q.dt;
        xi.x += vi.x * q.dt;
        xi.y += vi.y * q.dt;
        xi.z += vi.z * q.dt;
        v[i] = vi;
        x[i] = xi;
    }
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

// compute_sph_density_and_error_cuda + update_lamb_cuda (PBFCudaImpl.cu:70-122):
//   rho_i = max(sum_j m_j W_ij, rho0),  C_i = rho_i/rho0 - 1,
//   S_i = sum_j |gradW_ij/rho0|^2 + |sum_j gradW_ij / rho0|^2,
//   lambda_i = -C_i / (S_i + 1e-6)          (mass-free gradient, as the reference).
// Boundary lambda is defined as 0 (the reference never writes it).
template <bool MW, uint32_t L>
__global__ __launch_bounds__(256, VTP_PASS_MIN_WAVES) void k_pbf_lambda(const float4 *__restrict__ x,
                                                    const int32_t *__restrict__ mat,
                                                    const uint32_t *__restrict__ nbr,
                                                    const uint32_t *__restrict__ cnt, uint32_t n,
                                                    uint32_t stride, PbfK q, float *__restrict__ lam,
                                                    float *__restrict__ dens, float *__restrict__ errc,
                                                    float4 *__restrict__ sg, float4 *__restrict__ xl,
                                                    int static_out) {
    const uint32_t g = slot_index(), i = g / L, sub = g % L;   // L lanes per particle
    if (i >= n) return;
    if (mat[i] != VTP_MAT_FLUID) {
        // slots are permuted every tick, so non-fluid slots are written too:
        // by the first iteration of a tick, or the build that computes it
        // (static particles do not move and the outputs are the same
        // buffers), and by every iteration of a taped one
        if (sub != 0 || !static_out) return;
        pbf_lambda_store_static(i, x[i], lam, dens, errc, sg, xl);
        return;
    }
    const float4 xi = x[i];
    const uint32_t w = cnt[i];
    const uint32_t nn = w & ~kDeadBit;
    float rho = ((w & kDeadBit) || sub != 0) ? 0.0f : xi.w * q.k.w0;
    float gsx = 0, gsy = 0, gsz = 0, s2 = 0;
    for_each_neighbor_coop<L>(
        nbr, stride, i, nn, sub, [&](uint32_t j) { return ldg(x, j); },
        [&](const float4 &xj) { pbf_lambda_pair<MW>(xi, xj, q.k, rho, gsx, gsy, gsz, s2); });
    rho = lane_sum<L>(rho);
    gsx = lane_sum<L>(gsx);
    gsy = lane_sum<L>(gsy);
    gsz = lane_sum<L>(gsz);
    s2 = lane_sum<L>(s2);
    if (sub != 0) return;
    pbf_lambda_store(i, xi, rho, gsx, gsy, gsz, s2, q, lam, dens, errc, sg, xl);
}

// compute_dx_cuda + apply_dx_cuda (PBFCudaImpl.cu:124-170):
//   dx_i = (1/rho0) sum_j (lambda_i + lambda_j) gradW_ij,
//   x_i <- x_i + dx_i,  v_i <- v_i + dx_i / dt.
// Neighbours come from xl = (x, lambda) written by the lambda pass: one
// 16-byte gather per neighbour.  MW (extension) weights the pair by m_j, read
// from the position buffer.
template <bool MW, uint32_t L>
__global__ __launch_bounds__(256, VTP_PASS_MIN_WAVES) void k_pbf_dx(const float4 *__restrict__ x, const float4 *__restrict__ xl,
                                                float4 *__restrict__ xo, float4 *__restrict__ v,
                                                const int32_t *__restrict__ mat, const uint32_t *__restrict__ nbr,
                                                const uint32_t *__restrict__ cnt, uint32_t n, uint32_t stride,
                                                PbfK q, int copy_boundary) {
    const uint32_t g = slot_index(), i = g / L, sub = g % L;
    if (i >= n) return;
    if (mat[i] != VTP_MAT_FLUID) {
        // a taped tick chains distinct buffers, which need the static particles too
        if (copy_boundary && sub == 0) xo[i] = x[i];
        return;
    }
    const float4 xi = x[i];
    const float li = xl[i].w;
    const uint32_t nn = cnt[i] & ~kDeadBit;
    float ax = 0, ay = 0, az = 0;
    for_each_neighbor_coop<L>(
        nbr, stride, i, nn, sub,
        [&](uint32_t j) {
            float4 e = ldg(xl, j);
            if (MW) e.w = (li + e.w) * ldg_w(x, j);
            else e.w = li + e.w;
            return e;
        },
        [&](const float4 &xj) {
            float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
            float r2 = dx * dx + dy * dy + dz * dz;
            float s = cubic_grad_factor(r2, q.k) * xj.w;
            ax = fmaf(dx, s, ax);
            ay = fmaf(dy, s, ay);
            az = fmaf(dz, s, az);
        });
    ax = lane_sum<L>(ax) * q.inv_rho0;
    ay = lane_sum<L>(ay) * q.inv_rho0;
    az = lane_sum<L>(az) * q.inv_rho0;
    if (sub != 0) return;
    xo[i] = make_float4(xi.x + ax, xi.y + ay, xi.z + az, xi.w);
    float4 vi = v[i];
    vi.x += ax * q.inv_dt;
    vi.y += ay * q.inv_dt;
    vi.z += az * q.inv_dt;
    v[i] = vi;
}

// XSPH_cuda + apply_dx_cuda (PBFCudaImpl.cu:172-194, 267-281):
//   dv_i = k sum_j V_j (v_j - v_i) W_ij,  dx_i = dv_i dt,
//   x_i <- x_i + dx_i,  v_i <- v_i + dx_i / dt.
// Writes every slot of the output buffers (boundary slots are copied), so the
// state buffers X[0]/V[0] are complete after the pass.
template <uint32_t L>
__global__ __launch_bounds__(256, VTP_PASS_MIN_WAVES) void k_pbf_xsph(const float4 *__restrict__ x, float4 *__restrict__ xo,
                                                  const float4 *__restrict__ v, float4 *__restrict__ vo,
                                                  const int32_t *__restrict__ mat,
                                                  const uint32_t *__restrict__ nbr,
                                                  const uint32_t *__restrict__ cnt, uint32_t n,
                                                  uint32_t stride, PbfK q) {
    const uint32_t g = slot_index(), i = g / L, sub = g % L;
    if (i >= n) return;
    const float4 xi = x[i];
    const float4 vi = v[i];
    if (mat[i] != VTP_MAT_FLUID) {
        if (sub != 0) return;
        xo[i] = xi;
        vo[i] = vi;
        return;
    }
    const uint32_t nn = cnt[i] & ~kDeadBit;
    float ax = 0, ay = 0, az = 0;
    struct XV { float4 x, v; };
    for_each_neighbor_coop<L>(
        nbr, stride, i, nn, sub, [&](uint32_t j) { return XV{ldg(x, j), ldg(v, j)}; },
        [&](const XV &d) {
            float dx = xi.x - d.x.x, dy = xi.y - d.x.y, dz = xi.z - d.x.z;
            float r2 = dx * dx + dy * dy + dz * dz;
            float s = d.v.w * cubic_w(r2, q.k);
            ax = fmaf(d.v.x - vi.x, s, ax);
            ay = fmaf(d.v.y - vi.y, s, ay);
            az = fmaf(d.v.z - vi.z, s, az);
        });
    ax = lane_sum<L>(ax);
    ay = lane_sum<L>(ay);
    az = lane_sum<L>(az);
    if (sub != 0) return;
    float c = q.xsph_k * q.dt;
    float dx = ax * c, dy = ay * c, dz = az * c;
    xo[i] = make_float4(xi.x + dx, xi.y + dy, xi.z + dz, xi.w);
    vo[i] = make_float4(vi.x + dx * q.inv_dt, vi.y + dy * q.inv_dt, vi.z + dz * q.inv_dt, vi.w);
}

// init_cuda + compute_rigid_particle_volume_cuda (PBFCudaImpl.cu:34-68):
// fluid V = V0, m = rho0_f V0; boundary V = 1 / sum_{j same material} W_ij
// (self included), m = rho0_b V.
__global__ void k_pbf_init(float4 *__restrict__ x, float4 *__restrict__ v, const int32_t *__restrict__ mat,
                           const uint32_t *__restrict__ nbr, const uint32_t *__restrict__ cnt, uint32_t n,
                           uint32_t stride, Cubic k, float V0, float rho0_f, float rho0_b) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t mi = mat[i];
    float vol = V0, mass = rho0_f * V0;
    if (mi == VTP_MAT_BOUNDARY) {
        const float4 xi = x[i];
        const uint32_t w = cnt[i];
        const uint32_t nn = w & ~kDeadBit;
        float delta = (w & kDeadBit) ? 0.0f : k.w0;
    #pragma unroll 4
    for (uint32_t t = 0; t < nn; ++t) {
            uint32_t j = nbr_entry(nbr, stride, i, t);
            if (mat[j] != mi) continue;
            float4 xj = x[j];
            float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
            delta += cubic_w(dx * dx + dy * dy + dz * dz, k);
        }
        vol = 1.0f / delta;
        mass = rho0_b * vol;
    }
    x[i].w = mass;
    v[i].w = vol;
}

// ------------------------------------------------------------------ host side
void common_upload_and_search(Ctx &c);
void launch_step_begin(Ctx &c);
void common_reorder(Ctx &c, const float4 *xs, float4 *xd, float4 *xd2, const float4 *vs, float4 *vd,
                    float4 *vd2);

void pbf_initialize(Ctx &c) {
    const uint32_t n = c.n;
    c.lam.alloc(n);
    c.xl.alloc(n);
    c.dens.alloc(n);
    c.err_c.alloc(n);
    VTP_HIP(hipMemsetAsync(c.lam.p, 0, c.lam.bytes(), c.stream));
    VTP_HIP(hipMemsetAsync(c.dens.p, 0, c.dens.bytes(), c.stream));
    VTP_HIP(hipMemsetAsync(c.err_c.p, 0, c.err_c.bytes(), c.stream));
    common_upload_and_search(c);   // state in X[0]/V[0], lists built
    hipLaunchKernelGGL(k_pbf_init, dim3(div_up(n, 256)), dim3(256), 0, c.stream, c.X[0].p, c.V[0].p, c.mat.p,
                       c.ugns.nbr.p, c.ugns.cnt.p, n, c.ugns.stride, c.kern, c.V0, c.p.f_rest_density,
                       c.p.b_rest_density);
    VTP_HIP(hipGetLastError());
}

static void pbf_enqueue_step(Ctx &c, PbfTapeStep *tape = nullptr) {
    const uint32_t n = c.n;
    const PbfK q = pbf_constants(c);
    Ugns &u = c.ugns;
    const uint32_t nb = div_up(n, 256);
    launch_step_begin(c);
    c.timer.begin(c.stream, "pbf_predict_hash");
    hipLaunchKernelGGL(k_pbf_predict_hash, dim3(nb), dim3(256), 0, c.stream, c.X[0].p, c.V[0].p, c.mat.p,
                       c.scene.p, n, q, c.grid, u.key[0].p, u.val[0].p, c.status.p);
    c.timer.end(c.stream);
    // A taped tick writes its permutation, lists and per-iteration positions
    // straight into the tape: the search's buffers are lent to the tape's for
    // the tick, and the iterations chain the tape's position buffers.
    struct Lend {
        Ugns &u;
        PbfTapeStep *t;
        void swap() {
            if (!t) return;
            std::swap(u.perm.p, t->perm);
            std::swap(u.nbr.p, t->nbr);
            std::swap(u.cnt.p, t->cnt);
        }
        ~Lend() { swap(); }
    } lend{u, tape};
    lend.swap();
    u.sort_and_rank(c.orig.p, c.stream, c.timer);
    c.timer.begin(c.stream, "pbf_reorder");
    float4 *const x0 = tape ? tape->x[0] : c.X[1].p;
    common_reorder(c, c.X[0].p, x0, tape ? nullptr : c.X[2].p, c.V[0].p, c.V[1].p, nullptr);
    c.timer.end(c.stream);
    const uint32_t L = c.opt_lanes;
    // The build computes the first iteration's density + lambda pass while it
    // holds each listed neighbour (one lane per particle: the same sums in
    // the same order as the pass).
    const bool build_lambda = c.opt_build_lambda && L == 1 && c.p.iteration_num > 0;
    const PbfLambdaOut lo{q, c.mat.p, tape ? tape->lam[0] : c.lam.p, c.dens.p, tape ? tape->C[0] : c.err_c.p,
                          tape ? tape->sg[0] : nullptr, c.xl.p};
    // no kernel of a PBF tick (nor its adjoint) walks a static particle's list
    u.build_neighbors(x0, c.stream, c.timer, c.opt_static_lists ? nullptr : c.mat.p, build_lambda ? &lo : nullptr);
    if (tape) VTP_HIP(hipMemcpyAsync(tape->orig, c.orig.p, n * 4, hipMemcpyDeviceToDevice, c.stream));
    float4 *cur = x0, *nxt = c.X[2].p;
    const uint32_t bl = c.block("pbf_lambda"), bd = c.block("pbf_dx"), bx = c.block("pbf_xsph");
    for (uint32_t it = 0; it < c.p.iteration_num; ++it) {
        float *lam = tape ? tape->lam[it] : c.lam.p;
        float *errc = tape ? tape->C[it] : c.err_c.p;
        float4 *sg = tape ? tape->sg[it] : nullptr;
        if (tape) nxt = tape->x[it + 1];
        const dim3 gl(div_up(n * L, bl)), gd(div_up(n * L, bd));
#define VTP_LAMBDA(MW, LL)                                                                                     \
    hipLaunchKernelGGL((k_pbf_lambda<MW, LL>), gl, dim3(bl), 0, c.stream, cur, c.mat.p, u.nbr.p, u.cnt.p, n,    \
                       u.stride, q, lam, c.dens.p, errc, sg, c.xl.p, (tape || it == 0) ? 1 : 0)
#define VTP_DX(MW, LL)                                                                                         \
    hipLaunchKernelGGL((k_pbf_dx<MW, LL>), gd, dim3(bd), 0, c.stream, cur, c.xl.p, nxt, c.V[1].p, c.mat.p,    \
                       u.nbr.p, u.cnt.p, n, u.stride, q, tape ? 1 : 0)
        if (it > 0 || !build_lambda) {
            c.timer.begin(c.stream, "pbf_lambda");
            if (q.mass_weighted) {
                if (L == 4) VTP_LAMBDA(true, 4); else if (L == 2) VTP_LAMBDA(true, 2); else VTP_LAMBDA(true, 1);
            } else {
                if (L == 4) VTP_LAMBDA(false, 4); else if (L == 2) VTP_LAMBDA(false, 2); else VTP_LAMBDA(false, 1);
            }
            c.timer.end(c.stream);
        }
        c.timer.begin(c.stream, "pbf_dx");
        if (q.mass_weighted) {
            if (L == 4) VTP_DX(true, 4); else if (L == 2) VTP_DX(true, 2); else VTP_DX(true, 1);
        } else {
            if (L == 4) VTP_DX(false, 4); else if (L == 2) VTP_DX(false, 2); else VTP_DX(false, 1);
        }
#undef VTP_LAMBDA
#undef VTP_DX
        c.timer.end(c.stream);
        std::swap(cur, nxt);
    }
    if (tape) VTP_HIP(hipMemcpyAsync(tape->v_last, c.V[1].p, n * sizeof(float4), hipMemcpyDeviceToDevice, c.stream));
    c.timer.begin(c.stream, "pbf_xsph");
    if (L == 4)
        hipLaunchKernelGGL(k_pbf_xsph<4>, dim3(div_up(n * 4, bx)), dim3(bx), 0, c.stream, cur, c.X[0].p, c.V[1].p,
                           c.V[0].p, c.mat.p, u.nbr.p, u.cnt.p, n, u.stride, q);
    else
        hipLaunchKernelGGL(k_pbf_xsph<1>, dim3(div_up(n, bx)), dim3(bx), 0, c.stream, cur, c.X[0].p, c.V[1].p,
                           c.V[0].p, c.mat.p, u.nbr.p, u.cnt.p, n, u.stride, q);
    c.timer.end(c.stream);
}

// One tick that also records what the reverse pass needs (no graph).
void pbf_step_taped(Ctx &c, PbfTapeStep &tape) {
    pbf_enqueue_step(c, &tape);
    VTP_HIP(hipGetLastError());
}

void pbf_capture_reset(Ctx &c) {
    if (c.graph_exec) (void)hipGraphExecDestroy(c.graph_exec);
    if (c.graph) (void)hipGraphDestroy(c.graph);
    c.graph_exec = nullptr;
    c.graph = nullptr;
    ++c.capture_epoch;   // graphs captured elsewhere (the adjoint's) are stale as well
}

void pbf_step(Ctx &c) {
    c.lists_stale = false;
    if (c.opt_graph && !c.timer.enabled) {
        if (!c.graph_exec) {
            VTP_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
            pbf_enqueue_step(c);
            hipGraph_t g = nullptr;
            VTP_HIP(hipStreamEndCapture(c.stream, &g));
            c.graph = g;
            VTP_HIP(hipGraphInstantiate(&c.graph_exec, c.graph, nullptr, nullptr, 0));
        }
        VTP_HIP(hipGraphLaunch(c.graph_exec, c.stream));
    } else {
        pbf_enqueue_step(c);
    }
    VTP_HIP(hipGetLastError());
}

}  // namespace vtp
