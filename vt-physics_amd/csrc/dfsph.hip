// dfsph.hip - divergence-free SPH step (DFSPHCudaImpl.cu:38-620, DFSPHSolver.cpp:339-406)
// and the pressure solve shared with IMM / IMM-CT.
//
// Structure on the device:
//   * delta (density predictor) and kappa share one neighbour pass; the
//     residual is reduced on the device in FP64 per attach object (wavefront
//     sums, one atomic per wavefront), replacing the reference's full-array
//     D2H copy + host mean per iteration (CommonFunc.cpp:10-21);
//   * a one-thread "decide" kernel applies the reference's loop rule
//     (iterate while residual >= threshold and iter <= cap) and raises a
//     device flag; every kernel of the loop returns immediately once the flag
//     is up, so the host enqueues chunks of iterations and checks the flag
//     once per chunk instead of once per iteration;
//   * viscosity and surface normal share a pass; tension and acc->vel share
//     a pass; the final pressure application is fused with update_pos.
#include "vtp_internal.h"

namespace vtp {

struct DfK {
    Cubic k;
    float rho0_f, V0, vis0;
    float gamma, c_sigma, h6_64, h2_001;
    float gx, gy, gz;
    float r2;           // 2 * particle radius (CFL length)
    float cfl, dt_max;
};

static DfK df_constants(const Ctx &c) {
    DfK q;
    q.k = c.kern;
    q.rho0_f = c.p.f_rest_density;
    q.V0 = c.V0;
    q.vis0 = c.p.f_rest_viscosity;
    q.gamma = c.p.surface_tension_coefficient;
    const float pi = 3.14159265f;
    q.c_sigma = 32.0f / pi / std::pow(c.h, 9.0f);
    q.h6_64 = std::pow(c.h, 6.0f) / 64.0f;
    q.h2_001 = 0.001f * c.h * c.h;
    q.gx = c.p.gravity[0];
    q.gy = c.p.gravity[1];
    q.gz = c.p.gravity[2];
    q.r2 = 2.0f * c.p.particle_radius;
    q.cfl = c.p.cfl_number;
    q.dt_max = c.p.dt_max;
    return q;
}

struct Loop {   // device pointers of one pressure-solve loop
    uint32_t *status;
    SolverScalars *sc;
    const double *obj_w;       // per object: 1/(particles * fluid objects) for fluid objects, else 0
    double *partial;           // one residual partial sum per workgroup of the delta/kappa pass
    uint32_t nobj;
};

// ------------------------------------------------------------------ init
// init_cuda + update_mass_cuda + compute_rigid_volume_cuda (DFSPHCudaImpl.cu:38-83):
// fluid V = V0, m = rho0 V0, rest = rho0; boundary V = 1/sum W (same
// material), rest = V * rho0_b (the reference's mass-like quirk), mass 0
// (never set by the reference).
__global__ void k_df_init(float4 *__restrict__ x, float4 *__restrict__ v, float *__restrict__ rest,
                          float *__restrict__ vis, const int32_t *__restrict__ mat,
                          const uint32_t *__restrict__ nbr, const uint32_t *__restrict__ cnt, uint32_t n,
                          uint32_t stride, DfK q, float rho0_b) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t mi = mat[i];
    float V = q.V0, m = q.rho0_f * q.V0, r = q.rho0_f;
    if (mi == VTP_MAT_BOUNDARY) {
        const float4 xi = x[i];
        const uint32_t w = cnt[i];
        const uint32_t nn = w & ~kDeadBit;
        float d = (w & kDeadBit) ? 0.0f : q.k.w0;
        for (uint32_t t = 0; t < nn; ++t) {
            uint32_t j = nbr_entry(nbr, stride, i, t);
            if (mat[j] != mi) continue;
            float4 xj = x[j];
            float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
            d += cubic_w(dx * dx + dy * dy + dz * dz, q.k);
        }
        V = 1.0f / d;
        r = V * rho0_b;
        m = 0.0f;
    }
    x[i].w = m;
    v[i].w = V;
    rest[i] = r;
    vis[i] = q.vis0;
}

__global__ void k_scatter_attach_f32(const float *__restrict__ src, const uint32_t *__restrict__ orig, uint32_t n,
                                     float *__restrict__ dst) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[orig[i]] = src[i];
}

// --------------------------------------------------------------- CFL (extension)
__global__ void k_df_vmax(const float4 *__restrict__ v, const int32_t *__restrict__ mat, uint32_t n,
                          SolverScalars *sc) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    float s = 0.0f;
    if (i < n && mat[i] == VTP_MAT_FLUID) {
        float4 vi = v[i];
        s = sqrtf(vi.x * vi.x + vi.y * vi.y + vi.z * vi.z);
    }
    s = wave_max(s);
    if ((threadIdx.x & 63) == 0 && s > 0.0f) atomicMax(&sc->vmax_bits, __float_as_uint(s));
}

// dt = min(dt_max, cfl * 2r / max|v|); resets the running maximum.
__global__ void k_df_cfl_dt(SolverScalars *sc, DfK q, int enable) {
    if (threadIdx.x != 0) return;
    if (enable) {
        float vmax = __uint_as_float(sc->vmax_bits);
        float dt = q.dt_max;
        if (vmax > 0.0f) dt = fminf(q.dt_max, q.cfl * q.r2 / vmax);
        sc->dt = dt;
        sc->vmax = vmax;
    }
    sc->vmax_bits = 0;
}

// ------------------------------------------------------------------ precompute
// compute_compression_ratio_cuda + compute_df_beta_cuda (DFSPHCudaImpl.cu:85-140)
// and v* = v.  alpha is consumed only for fluid particles; boundary alpha = 0.
__global__ __launch_bounds__(256) void k_df_precompute(const float4 *__restrict__ x, const float4 *__restrict__ v,
                                                       float4 *__restrict__ vadv,
                                                       const int32_t *__restrict__ mat,
                                                       const uint32_t *__restrict__ nbr,
                                                       const uint32_t *__restrict__ cnt, uint32_t n,
                                                       uint32_t stride, DfK q, float *__restrict__ cr,
                                                       float *__restrict__ alpha, float4 *__restrict__ g1,
                                                       const uint32_t *__restrict__ obj_of,
                                                       float4 *__restrict__ sconst) {
    uint32_t i = slot_index();
    if (i >= n) return;
    const float4 vi = v[i];
    const float4 xi = x[i];
    vadv[i] = vi;
    if (mat[i] != VTP_MAT_FLUID) {
        cr[i] = 0.0f;
        alpha[i] = 0.0f;
        g1[i] = make_float4(0, 0, 0, 0);
        sconst[i] = make_float4(0.0f, 0.0f, vi.w, 0.0f);   // alpha = 0: not fluid
        return;
    }
    const uint32_t w = cnt[i];
    const uint32_t nn = w & ~kDeadBit;
    float c = (w & kDeadBit) ? 0.0f : vi.w * q.k.w0;
    float a1x = 0, a1y = 0, a1z = 0, a2 = 1e-6f;
    // A neighbour is fluid exactly when its mass is not zero (fluid m =
    // rho0 V0 > 0, boundary m = 0, as k_df_init sets them), so the material
    // needs no gather of its own.
    struct D { float4 x; float V; };
    for_each_neighbor(
        nbr, stride, i, nn, [&](uint32_t j) { return D{ldg(x, j), ldg_w(v, j)}; },
        [&](const D &d) {
            float dx = xi.x - d.x.x, dy = xi.y - d.x.y, dz = xi.z - d.x.z;
            float W, f;
            cubic_w_grad(dx * dx + dy * dy + dz * dz, q.k, W, f);
            c = fmaf(d.V, W, c);
            float gx = dx * f, gy = dy * f, gz = dz * f;
            a1x = fmaf(d.V, gx, a1x);
            a1y = fmaf(d.V, gy, a1y);
            a1z = fmaf(d.V, gz, a1z);
            if (d.x.w != 0.0f) a2 += (gx * gx + gy * gy + gz * gz) * d.V * d.V / d.x.w;
        });
    float a = (a1x * a1x + a1y * a1y + a1z * a1z) / xi.w + a2;
    cr[i] = c;
    alpha[i] = fmaxf(a, 1e-6f);
    sconst[i] = make_float4(c, fmaxf(a, 1e-6f), vi.w, __uint_as_float(obj_of[i]));
    // sum_j V_j gradW_ij, reused by every v* update of the tick, and V_i / m_i
    // (w != 0 marks a fluid slot for that pass)
    g1[i] = make_float4(a1x, a1y, a1z, vi.w / xi.w);
}

// ------------------------------------------------------------------ pressure solve
__global__ void k_solve_begin(uint32_t *status, SolverScalars *sc, uint32_t nobj) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        status[kStDone] = 0;
        status[kStIter] = 0;
    }
    (void)sc;
    (void)nobj;
}

// compute_delta_compression_ratio + update_delta_compression_ratio_from_vel_adv
// + compute_kappa_*_from_delta_compression_ratio (DFSPHCudaImpl.cu:142-196, 355-367):
//   delta_i = max(0, cr_i - 1 + dt sum_j V_j (v*_i - v*_j) . gradW_ij),
//   kappa_i = delta_i / alpha_i / dt^2 / V_i  (boundary kappa = 0),
// with the per-object sum of delta added to sc->obj_sum in FP64.
// sconst[i] = (cr, alpha, V, object id) from the precompute pass: constant
// through the solves of a tick, one 16-byte load.  alpha is exactly 0 for a
// static slot and at least 1e-6 (or NaN, which the residual then carries)
// for a fluid one.  L lanes per particle share its list (L = 1: one lane).
template <uint32_t L>
__global__ __launch_bounds__(256) void k_df_delta_kappa(const float4 *__restrict__ x,
                                                        const float4 *__restrict__ vadv,
                                                        const float4 *__restrict__ sconst,
                                                        const uint32_t *__restrict__ nbr,
                                                        const uint32_t *__restrict__ cnt, uint32_t n,
                                                        uint32_t stride, Cubic k, float *__restrict__ delta,
                                                        float *__restrict__ kappa, float4 *__restrict__ xk, Loop lp,
                                                        int static_out) {
    if (lp.status[kStDone]) return;
    const uint32_t g = slot_index(), i = g / L, sub = g % L;
    const float dt = lp.sc->dt;
    double contrib = 0.0;
    if (i < n) {
        const float4 sc = sconst[i];
        if (sc.y != 0.0f) {
            const float4 xi = x[i];
            const float4 ai = vadv[i];
            const uint32_t nn = cnt[i] & ~kDeadBit;
            float d = 0.0f;
            struct D { float4 x, a; };
            for_each_neighbor_coop<L>(
                nbr, stride, i, nn, sub,
                [&](uint32_t j) { return D{ldg(x, j), ldg(vadv, j)}; },   // vadv.w = V_j
                [&](const D &e) {
                    float dx = xi.x - e.x.x, dy = xi.y - e.x.y, dz = xi.z - e.x.z;
                    float f = cubic_grad_factor(dx * dx + dy * dy + dz * dz, k);
                    float dot = (dx * (ai.x - e.a.x) + dy * (ai.y - e.a.y) + dz * (ai.z - e.a.z)) * f;
                    d = fmaf(dot, e.a.w, d);
                });
            d = lane_sum<L>(d);
            if (sub == 0) {
                float dl = fmaxf(sc.x - 1.0f + d * dt, 0.0f);
                float inv_dt = 1.0f / dt;
                const float Vi = sc.z;
                const float ki = dl / sc.y * (inv_dt * inv_dt) / Vi;
                delta[i] = dl;
                kappa[i] = ki;
                xk[i] = make_float4(xi.x, xi.y, xi.z, Vi * ki);
                contrib = (double)dl * lp.obj_w[__float_as_uint(sc.w)];
            }
        } else if (static_out && sub == 0) {
            // static particles do not move during a solve: their kappa and
            // (x, 0) records are written by its first iteration only
            kappa[i] = 0.0f;
            const float4 xi = x[i];
            xk[i] = make_float4(xi.x, xi.y, xi.z, 0.0f);
        }
    }
    // residual = sum_i delta_i w_obj(i) (the mean over fluid objects of the
    // per-object mean of delta): one FP64 partial per workgroup, reduced in a
    // fixed order by k_df_decide, so the residual is deterministic and no
    // atomic touches a shared address
    __shared__ double s_part[kMaxWavesPerBlock];
    const double wsum = wave_sum(contrib);
    const uint32_t w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if ((threadIdx.x & 63) == 0) s_part[w] = wsum;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (uint32_t q = 0; q < nw; ++q) t += s_part[q];
        lp.partial[blockIdx.x] = t;
    }
}

// vf_update_vel_adv_from_kappa_* (DFSPHCudaImpl.cu:198-220, 369-391):
//   v*_i -= dt V_i V_j / m_i (kappa_i + kappa_j) gradW_ij over all neighbours,
// evaluated as dt V_i / m_i (kappa_i g1_i + sum_j (V_j kappa_j) gradW_ij) with
// g1_i = sum_j V_j gradW_ij from the precompute pass (positions are fixed
// during the solve): a neighbour is one 16-byte (x_j, V_j kappa_j) gather.
template <uint32_t L>
__global__ __launch_bounds__(256) void k_df_update_vadv(const float4 *__restrict__ xk,
                                                        const float4 *__restrict__ g1,
                                                        float4 *__restrict__ vadv,
                                                        const uint32_t *__restrict__ nbr,
                                                        const uint32_t *__restrict__ cnt, uint32_t n,
                                                        uint32_t stride, Cubic k, const float *__restrict__ kappa,
                                                        Loop lp) {
    if (lp.status[kStDone]) return;
    const uint32_t g = slot_index(), i = g / L, sub = g % L;
    if (i >= n) return;
    // g1[i] = (sum_j V_j gradW_ij, V_i / m_i), w = 0 for a static slot; the
    // particle's own (x, V kappa) record, written by the delta pass, gives x_i
    const float4 gi = g1[i];
    if (gi.w == 0.0f) return;
    const float dt = lp.sc->dt;
    const float4 xi = xk[i];
    const float ki = kappa[i];
    const uint32_t nn = cnt[i] & ~kDeadBit;
    float ax = 0, ay = 0, az = 0;
    for_each_neighbor_coop<L>(
        nbr, stride, i, nn, sub, [&](uint32_t j) { return ldg(xk, j); },   // (x_j, V_j kappa_j)
        [&](const float4 &e) {
            float dx = xi.x - e.x, dy = xi.y - e.y, dz = xi.z - e.z;
            float s = cubic_grad_factor(dx * dx + dy * dy + dz * dz, k) * e.w;
            ax = fmaf(dx, s, ax);
            ay = fmaf(dy, s, ay);
            az = fmaf(dz, s, az);
        });
    ax = lane_sum<L>(ax);
    ay = lane_sum<L>(ay);
    az = lane_sum<L>(az);
    if (sub != 0) return;
    ax = fmaf(ki, gi.x, ax);
    ay = fmaf(ki, gi.y, ay);
    az = fmaf(ki, gi.z, az);
    float c = dt * gi.w;
    float4 a = vadv[i];
    a.x -= c * ax;
    a.y -= c * ay;
    a.z -= c * az;
    vadv[i] = a;
}

// Loop rule of vfsph_div / vfsph_incomp (DFSPHCudaImpl.cu:458-502): after the
// v* update, stop when the residual of this iteration is below the threshold
// or more than `cap` iterations ran (then the crash bit is raised).  One
// workgroup sums the per-workgroup partials of the delta/kappa pass in a
// fixed order.
#ifndef VTP_DECIDE_THREADS
#define VTP_DECIDE_THREADS 1024
#endif
// One workgroup of kDecideThreads: every thread sums a strided share of the
// partials, each wave reduces its lanes with shuffles and one barrier joins
// the waves' sums, so the reduction order is fixed for a given partial count.
constexpr uint32_t kDecideThreads = VTP_DECIDE_THREADS;
__global__ __launch_bounds__(kDecideThreads) void k_df_decide(Loop lp, uint32_t nparts, float thr, uint32_t cap,
                                                               uint32_t iter_word, uint32_t crash_bit, int which) {
    if (lp.status[kStDone]) return;
    __shared__ double s_red[kDecideThreads / 64];
    double t = 0.0;
#pragma unroll 4
    for (uint32_t k = threadIdx.x; k < nparts; k += kDecideThreads) t += lp.partial[k];
    t = wave_sum(t);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x != 0) return;
    double res = 0.0;
#pragma unroll
    for (uint32_t w = 0; w < kDecideThreads / 64; ++w) res += s_red[w];
    const uint32_t iter = lp.status[kStIter] + 1;
    lp.status[kStIter] = iter;
    if (which == 0) lp.sc->div_residual = res; else lp.sc->inc_residual = res;
    if (res < (double)thr || iter > cap) {
        lp.status[kStDone] = 1;
        lp.status[iter_word] = iter;
        if (iter > cap) atomicOr(&lp.status[kStCrash], crash_bit);
    }
}

// apply_pressure_acc (DFSPHCudaImpl.cu:505-521): acc = (v* - v)/dt,
// v += acc dt, v* = v.  With `xo` set it also performs update_pos (:393-404)
// and writes the complete state (boundary slots copied) to xo / vo.
__global__ void k_df_apply_pressure(const float4 *__restrict__ x, float4 *__restrict__ v,
                                    float4 *__restrict__ vadv, float4 *__restrict__ acc,
                                    const int32_t *__restrict__ mat, uint32_t n, const SolverScalars *sc,
                                    float4 *__restrict__ xo, float4 *__restrict__ vo) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float dt = sc->dt;
    float4 vi = v[i];
    if (mat[i] != VTP_MAT_FLUID) {
        acc[i] = make_float4(0, 0, 0, 0);
        if (xo) {
            xo[i] = x[i];
            vo[i] = vi;
        }
        return;
    }
    const float inv_dt = 1.0f / dt;
    float4 a = vadv[i];
    float4 ac = make_float4((a.x - vi.x) * inv_dt, (a.y - vi.y) * inv_dt, (a.z - vi.z) * inv_dt, 0.0f);
    vi.x += ac.x * dt;
    vi.y += ac.y * dt;
    vi.z += ac.z * dt;
    acc[i] = ac;
    vadv[i] = make_float4(vi.x, vi.y, vi.z, a.w);
    if (xo) {
        float4 xi = x[i];
        xo[i] = make_float4(xi.x + vi.x * dt, xi.y + vi.y * dt, xi.z + vi.z * dt, xi.w);
        vo[i] = vi;
    } else {
        v[i] = vi;
    }
}

// Gravity + explicit Laplacian viscosity (DFSPHCudaImpl.cu:253-296) and the
// surface normal (:298-321), one pass over same-material neighbours:
//   acc_i = g + 10 sum_j (mu_i+mu_j)/2 m_j/rho0_j (v_ij.x_ij)/(|x_ij|^2 + 0.001 h^2) gradW_ij,
//   n_i = h sum_j m_j/rho0_j gradW_ij.
__global__ __launch_bounds__(256) void k_df_vis_normal(const float4 *__restrict__ x, const float4 *__restrict__ v,
                                                       const int32_t *__restrict__ mat,
                                                       const uint32_t *__restrict__ nbr,
                                                       const uint32_t *__restrict__ cnt, uint32_t n,
                                                       uint32_t stride, DfK q, float4 *__restrict__ acc,
                                                       float4 *__restrict__ normal) {
    uint32_t i = slot_index();
    if (i >= n) return;
    const int32_t mi = mat[i];
    if (mi != VTP_MAT_FLUID) {
        acc[i] = make_float4(0, 0, 0, 0);
        normal[i] = make_float4(0, 0, 0, 0);   // slots move between ticks: clear what a fluid particle left
        return;
    }
    const float4 xi = x[i];
    const float4 vi = v[i];
    const uint32_t nn = cnt[i] & ~kDeadBit;
    float ax = q.gx, ay = q.gy, az = q.gz;
    float nx = 0, ny = 0, nz = 0;
    // Same-material neighbours of a fluid particle are the fluid ones, those
    // with non-zero mass (boundary m = 0), whose rest density is rho0_f and
    // whose viscosity is the solver's one fPartRestViscosity (k_df_init sets
    // every particle's, and nothing changes them), so a neighbour is its
    // (x, m) and (v, V) records and mu_i is that viscosity too.
    const float visi = q.vis0;
    const float vis_ij = 0.5f * (visi + visi);   // (mu_i + mu_j) / 2 with mu_j = mu_i
    struct D { float4 x, v; };
    for_each_neighbor(
        nbr, stride, i, nn, [&](uint32_t j) { return D{ldg(x, j), ldg(v, j)}; },
        [&](const D &d) {
            if (d.x.w == 0.0f) return;
            const float mr = d.x.w / q.rho0_f;
            float dx = xi.x - d.x.x, dy = xi.y - d.x.y, dz = xi.z - d.x.z;
            float r2 = dx * dx + dy * dy + dz * dz;
            float f = cubic_grad_factor(r2, q.k);
            float vdx = (vi.x - d.v.x) * dx + (vi.y - d.v.y) * dy + (vi.z - d.v.z) * dz;
            float pi = vis_ij * mr * vdx / (r2 + q.h2_001);
            float s = 10.0f * pi * f;
            ax = fmaf(dx, s, ax);
            ay = fmaf(dy, s, ay);
            az = fmaf(dz, s, az);
            float sn = q.k.h * mr * f;
            nx = fmaf(dx, sn, nx);
            ny = fmaf(dy, sn, ny);
            nz = fmaf(dz, sn, nz);
        });
    acc[i] = make_float4(ax, ay, az, 0.0f);
    normal[i] = make_float4(nx, ny, nz, 0.0f);
}

// Surface tension (DFSPHCudaImpl.cu:323-353) + acc_2_vel (:241-251):
//   acc_i += sum_{j same mat} k_ij (-gamma m_i m_j C(r) x_ij/r - gamma m_i (n_i - n_j)),
//   k_ij = 2 rho0_i / (rho0_i + rho0_j);  v += acc dt;  v* = v.
__global__ __launch_bounds__(256) void k_df_tension_apply(const float4 *__restrict__ x, float4 *__restrict__ v,
                                                          float4 *__restrict__ vadv,
                                                          const int32_t *__restrict__ mat,
                                                          const float4 *__restrict__ normal,
                                                          const uint32_t *__restrict__ nbr,
                                                          const uint32_t *__restrict__ cnt, uint32_t n,
                                                          uint32_t stride, DfK q, float4 *__restrict__ acc,
                                                          const SolverScalars *sc) {
    uint32_t i = slot_index();
    if (i >= n) return;
    const int32_t mi = mat[i];
    if (mi != VTP_MAT_FLUID) return;
    float4 a = acc[i];
    if (q.gamma != 0.0f) {
        const float4 xi = x[i];
        const float4 ni = normal[i];
        const float ri = q.rho0_f;   // a fluid particle's rest density (k_df_init)
        const uint32_t nn = cnt[i] & ~kDeadBit;
        float sx = 0, sy = 0, sz = 0;
        // same-material neighbours: the fluid ones (non-zero mass), whose
        // rest density is rho0_f (as in k_df_vis_normal)
        struct D { float4 x, n; };
        for_each_neighbor(
            nbr, stride, i, nn, [&](uint32_t j) { return D{ldg(x, j), ldg(normal, j)}; },
            [&](const D &d) {
                if (d.x.w == 0.0f) return;
                float dx = xi.x - d.x.x, dy = xi.y - d.x.y, dz = xi.z - d.x.z;
                float r = sqrtf(dx * dx + dy * dy + dz * dz);
                float kk = 2.0f * ri / (ri + q.rho0_f);
                float c1 = -q.gamma * xi.w * d.x.w * surface_tension_c(r, q.k.h, q.c_sigma, q.h6_64) / r;
                float c2 = -q.gamma * xi.w;
                sx += kk * (c1 * dx + c2 * (ni.x - d.n.x));
                sy += kk * (c1 * dy + c2 * (ni.y - d.n.y));
                sz += kk * (c1 * dz + c2 * (ni.z - d.n.z));
            });
        a.x += sx;
        a.y += sy;
        a.z += sz;
        acc[i] = a;
    }
    const float dt = sc->dt;
    float4 vi = v[i];
    vi.x += a.x * dt;
    vi.y += a.y * dt;
    vi.z += a.z * dt;
    v[i] = vi;
    vadv[i] = vi;
}

// ------------------------------------------------------------------ host side
void common_upload_and_search(Ctx &c);
void common_reorder(Ctx &c, const float4 *xs, float4 *xd, float4 *xd2, const float4 *vs, float4 *vd, float4 *vd2);
void launch_step_begin(Ctx &c);
void launch_hash(Ctx &c, const float4 *x);

static Loop make_loop(Ctx &c) {
    return Loop{c.status.p, c.scal.p, c.obj_w.p, c.partial.p, (uint32_t)c.objects.size()};
}

// Pressure solve with device-side termination; the host checks the flag once
// per chunk of enqueued iterations.  which: 0 = divergence, 1 = density.
void df_pressure_solve(Ctx &c, const float4 *x, int which) {
    const uint32_t n = c.n;
    Ugns &u = c.ugns;
    Loop lp = make_loop(c);
    const float thr = which == 0 ? c.p.div_free_threshold : c.p.incomp_threshold;
    const uint32_t word = which == 0 ? kStDivIter : kStIncIter;
    const uint32_t bit = which == 0 ? VTP_CRASH_DIV_ITER_CAP : VTP_CRASH_INCOMP_ITER_CAP;
    const uint32_t cap = c.p.max_solver_iterations;
    const char *name_dk = which == 0 ? "df_div_delta_kappa" : "df_inc_delta_kappa";
    const char *name_up = which == 0 ? "df_div_update_vadv" : "df_inc_update_vadv";
    const uint32_t bd = c.block("df_delta_kappa"), bu = c.block("df_update_vadv");
    const uint32_t L = c.opt_lanes;   // lanes per particle in both solve passes
    // kappa_div and kappa_incomp are separate arrays, as in the reference
    // (DFSPHCudaImpl.cu:182-196 and 355-367), so both stay readable
    float *kap = which == 0 ? c.kappa_div.p : c.kappa.p;
    hipLaunchKernelGGL(k_solve_begin, dim3(1), dim3(256), 0, c.stream, c.status.p, c.scal.p, lp.nobj);
    uint32_t &hint = which == 0 ? c.iter_hint_div : c.iter_hint_inc;
    // The first chunk is the previous solve's iteration count; past it the
    // host checks every max(1, hint/4) iterations, which bounds the no-op
    // launches enqueued after the exit.  The stage timer drops the events of
    // those no-op iterations.
    const uint32_t step = c.opt_chunk ? c.opt_chunk : std::max(1u, hint / 4);
    uint32_t chunk = c.opt_chunk ? c.opt_chunk : std::max(1u, hint);
    uint32_t launched = 0;
    const size_t tmark = c.timer.mark();
    for (;;) {
        for (uint32_t it = 0; it < chunk && launched <= cap; ++it, ++launched) {
            c.timer.begin(c.stream, name_dk);
            const int so = launched == 0 ? 1 : 0;
            if (L == 4)
                hipLaunchKernelGGL(k_df_delta_kappa<4>, dim3(div_up(4 * n, bd)), dim3(bd), 0, c.stream, x, c.vadv.p,
                                   c.sconst.p, u.nbr.p, u.cnt.p, n, u.stride, c.kern, c.delta.p, kap, c.xk.p, lp, so);
            else if (L == 2)
                hipLaunchKernelGGL(k_df_delta_kappa<2>, dim3(div_up(2 * n, bd)), dim3(bd), 0, c.stream, x, c.vadv.p,
                                   c.sconst.p, u.nbr.p, u.cnt.p, n, u.stride, c.kern, c.delta.p, kap, c.xk.p, lp, so);
            else
                hipLaunchKernelGGL(k_df_delta_kappa<1>, dim3(div_up(n, bd)), dim3(bd), 0, c.stream, x, c.vadv.p,
                                   c.sconst.p, u.nbr.p, u.cnt.p, n, u.stride, c.kern, c.delta.p, kap, c.xk.p, lp, so);
            c.timer.end(c.stream);
            c.timer.begin(c.stream, name_up);
            if (L == 4)
                hipLaunchKernelGGL(k_df_update_vadv<4>, dim3(div_up(4 * n, bu)), dim3(bu), 0, c.stream, c.xk.p, c.g1.p,
                                   c.vadv.p, u.nbr.p, u.cnt.p, n, u.stride, c.kern, kap, lp);
            else if (L == 2)
                hipLaunchKernelGGL(k_df_update_vadv<2>, dim3(div_up(2 * n, bu)), dim3(bu), 0, c.stream, c.xk.p, c.g1.p,
                                   c.vadv.p, u.nbr.p, u.cnt.p, n, u.stride, c.kern, kap, lp);
            else
                hipLaunchKernelGGL(k_df_update_vadv<1>, dim3(div_up(n, bu)), dim3(bu), 0, c.stream, c.xk.p, c.g1.p,
                                   c.vadv.p, u.nbr.p, u.cnt.p, n, u.stride, c.kern, kap, lp);
            c.timer.end(c.stream);
            hipLaunchKernelGGL(k_df_decide, dim3(1), dim3(kDecideThreads), 0, c.stream, lp, div_up(L * n, bd), thr,
                               cap, word, bit, which);
        }
        read_status(c);
        if (c.h_status[kStDone] || launched > cap) break;
        chunk = step;
    }
    uint32_t iters = c.h_status[word];
    if (c.timer.enabled) c.timer.drop_from(tmark + 2 * (size_t)iters);   // two timed launches per iteration
    hint = iters;
    if (which == 0) c.total_div += iters; else c.total_inc += iters;
}

void dfsph_allocate_common(Ctx &c) {
    const uint32_t n = c.n;
    c.rest_density.alloc(n);
    c.vis.alloc(n);
    c.rest_attach.alloc(n);
    c.vis_attach.alloc(n);
    c.cr.alloc(n);
    c.alpha.alloc(n);
    c.sconst.alloc(n);
    c.delta.alloc(n);
    c.kappa_div.alloc(n);
    c.kappa.alloc(n);
    c.vadv.alloc(n);
    c.acc.alloc(n);
    c.normal.alloc(n);
    c.xk.alloc(n);
    c.g1.alloc(n);
    VTP_HIP(hipMemsetAsync(c.cr.p, 0, c.cr.bytes(), c.stream));
    VTP_HIP(hipMemsetAsync(c.alpha.p, 0, c.alpha.bytes(), c.stream));
    VTP_HIP(hipMemsetAsync(c.delta.p, 0, c.delta.bytes(), c.stream));
    VTP_HIP(hipMemsetAsync(c.kappa_div.p, 0, c.kappa_div.bytes(), c.stream));
    VTP_HIP(hipMemsetAsync(c.kappa.p, 0, c.kappa.bytes(), c.stream));
    VTP_HIP(hipMemsetAsync(c.acc.p, 0, c.acc.bytes(), c.stream));
    VTP_HIP(hipMemsetAsync(c.normal.p, 0, c.normal.bytes(), c.stream));
    // object tables for the residual mean
    const uint32_t no = (uint32_t)c.objects.size();
    std::vector<uint32_t> oc(no);
    std::vector<int32_t> of(no);
    for (uint32_t o = 0; o < no; ++o) {
        oc[o] = c.objects[o].end - c.objects[o].start;
        of[o] = c.objects[o].mat == VTP_MAT_FLUID ? 1 : 0;
    }
    c.obj_cnt.alloc(no);
    c.obj_fluid.alloc(no);
    VTP_HIP(hipMemcpyAsync(c.obj_cnt.p, oc.data(), no * 4, hipMemcpyHostToDevice, c.stream));
    VTP_HIP(hipMemcpyAsync(c.obj_fluid.p, of.data(), no * 4, hipMemcpyHostToDevice, c.stream));
    // residual weights (CommonFunc.cpp:10-21: mean over fluid objects of each object's mean)
    uint32_t nfo = 0;
    for (uint32_t o = 0; o < no; ++o) nfo += of[o];
    std::vector<double> ow(no, 0.0);
    for (uint32_t o = 0; o < no; ++o)
        if (of[o] && oc[o]) ow[o] = 1.0 / ((double)oc[o] * (double)nfo);
    c.obj_w.alloc(no);
    VTP_HIP(hipMemcpyAsync(c.obj_w.p, ow.data(), no * sizeof(double), hipMemcpyHostToDevice, c.stream));
    c.partial.alloc(div_up(4 * n, 64));   // 4 lanes per particle in the smallest workgroup of the delta/kappa pass
    SolverScalars s{};
    s.dt = c.p.time_step;
    VTP_HIP(hipMemcpyAsync(c.scal.p, &s, sizeof s, hipMemcpyHostToDevice, c.stream));
    VTP_HIP(hipStreamSynchronize(c.stream));
}

void dfsph_initialize(Ctx &c) {
    const uint32_t n = c.n;
    dfsph_allocate_common(c);
    common_upload_and_search(c);
    const DfK q = df_constants(c);
    hipLaunchKernelGGL(k_df_init, dim3(div_up(n, 256)), dim3(256), 0, c.stream, c.X[0].p, c.V[0].p,
                       c.rest_density.p, c.vis.p, c.mat.p, c.ugns.nbr.p, c.ugns.cnt.p, n, c.ugns.stride, q,
                       c.p.b_rest_density);
    hipLaunchKernelGGL(k_scatter_attach_f32, dim3(div_up(n, 256)), dim3(256), 0, c.stream, c.rest_density.p,
                       c.orig.p, n, c.rest_attach.p);
    hipLaunchKernelGGL(k_scatter_attach_f32, dim3(div_up(n, 256)), dim3(256), 0, c.stream, c.vis.p, c.orig.p, n,
                       c.vis_attach.p);
    VTP_HIP(hipMemcpyAsync(c.vadv.p, c.V[0].p, n * sizeof(float4), hipMemcpyDeviceToDevice, c.stream));
    VTP_HIP(hipGetLastError());
    VTP_HIP(hipStreamSynchronize(c.stream));
}

// DFSPHSolver::tick (DFSPHSolver.cpp:339-406).
void dfsph_step(Ctx &c) {
    const uint32_t n = c.n;
    const DfK q = df_constants(c);
    Ugns &u = c.ugns;
    const uint32_t nb = div_up(n, 256);
    launch_step_begin(c);
    if (c.p.cfl_enable) {
        hipLaunchKernelGGL(k_df_vmax, dim3(nb), dim3(256), 0, c.stream, c.V[0].p, c.mat.p, n, c.scal.p);
        hipLaunchKernelGGL(k_df_cfl_dt, dim3(1), dim3(64), 0, c.stream, c.scal.p, q, 1);
    }
    // neighbour search on the current positions; state moves to X[1]/V[1]
    c.timer.begin(c.stream, "df_hash");
    launch_hash(c, c.X[0].p);
    c.timer.end(c.stream);
    u.sort_and_rank(c.orig.p, c.stream, c.timer);
    c.timer.begin(c.stream, "df_reorder");
    common_reorder(c, c.X[0].p, c.X[1].p, nullptr, c.V[0].p, c.V[1].p, nullptr);
    c.timer.end(c.stream);
    // no tick kernel reads a static particle's list (k_df_init, at
    // initialize, is the only one): fluid lists only, unless asked for all
    u.build_neighbors(c.X[1].p, c.stream, c.timer, c.opt_static_lists ? nullptr : c.mat.p);
    float4 *x = c.X[1].p, *v = c.V[1].p;
    const uint32_t bp = c.block("df_precompute");
    c.timer.begin(c.stream, "df_precompute");
    hipLaunchKernelGGL(k_df_precompute, dim3(div_up(n, bp)), dim3(bp), 0, c.stream, x, v, c.vadv.p, c.mat.p, u.nbr.p,
                       u.cnt.p, n, u.stride, q, c.cr.p, c.alpha.p, c.g1.p, c.obj_of.p, c.sconst.p);
    c.timer.end(c.stream);
    df_pressure_solve(c, x, 0);
    c.timer.begin(c.stream, "df_apply_pressure");
    hipLaunchKernelGGL(k_df_apply_pressure, dim3(nb), dim3(256), 0, c.stream, x, v, c.vadv.p, c.acc.p, c.mat.p, n,
                       c.scal.p, nullptr, nullptr);
    c.timer.end(c.stream);
    const uint32_t bv = c.block("df_vis_normal");
    c.timer.begin(c.stream, "df_vis_normal");
    hipLaunchKernelGGL(k_df_vis_normal, dim3(div_up(n, bv)), dim3(bv), 0, c.stream, x, v, c.mat.p, u.nbr.p,
                       u.cnt.p, n, u.stride, q, c.acc.p, c.normal.p);
    c.timer.end(c.stream);
    c.timer.begin(c.stream, "df_tension_apply");
    hipLaunchKernelGGL(k_df_tension_apply, dim3(div_up(n, bv)), dim3(bv), 0, c.stream, x, v, c.vadv.p, c.mat.p,
                       c.normal.p, u.nbr.p, u.cnt.p, n, u.stride, q, c.acc.p, c.scal.p);
    c.timer.end(c.stream);
    df_pressure_solve(c, x, 1);
    c.timer.begin(c.stream, "df_apply_pressure_update_pos");
    hipLaunchKernelGGL(k_df_apply_pressure, dim3(nb), dim3(256), 0, c.stream, x, v, c.vadv.p, c.acc.p, c.mat.p, n,
                       c.scal.p, c.X[0].p, c.V[0].p);
    c.timer.end(c.stream);
    VTP_HIP(hipGetLastError());
    if (c.p.cfl_enable) {
        VTP_HIP(hipMemcpyAsync(c.h_scal, c.scal.p, sizeof(SolverScalars), hipMemcpyDeviceToHost, c.stream));
        VTP_HIP(hipStreamSynchronize(c.stream));
        c.sim_time += c.h_scal->dt;
    }
}

}  // namespace vtp
