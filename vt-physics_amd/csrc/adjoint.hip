// adjoint.hip - reverse mode of the PBF tick on the GPU (no reference counterpart).
//
// Rollout forward: ticks run as usual (graph replay allowed) and the full
// state (positions+mass, velocities+volume, slot->attach map) is saved every
// `ckpt_every` ticks.  Rollout backward: segments are processed last to
// first; each is recomputed from its checkpoint with a taped forward (lists,
// slot permutation, positions entering every constraint iteration, lambda,
// C, sum_j gradW and S, velocities entering XSPH) and then reversed tick by
// tick:
//   XSPH       x' = x + dt dv, v' = v + dv, dv_i = k sum_j V_j (v_j - v_i) W_ij
//   iteration  x' = x + dx, v' = v + dx/dt, dx_i = (1/rho0) sum_j (l_i + l_j) gradW_ij,
//              l_i = -C_i/(S_i + 1e-6), C_i = max(rho_i, rho0)/rho0 - 1,
//              S_i = (1/rho0^2)(sum_j |gradW_ij|^2 + |sum_j gradW_ij|^2)
//   reorder    x_new[i] = x_old[perm[i]]  ->  xbar_old[perm[i]] = xbar_new[i]
//   predict    v1 = v0 + g dt, x1 = x0 + v1 dt
// with neighbour lists held fixed, H_ij = f I + (f'(r)/r) r r^T the Hessian
// of W, and symmetric lists (j in N(i) <=> i in N(j)), which holds whenever
// no list is cut by maxNeighborNum.  The forward counts the particles whose
// list reached the cap in every build of the rollout (kStCapSeen); when any
// did, the rollout is refused with VTP_ENBRCAP and no gradient is returned.
#include <cstring>
#include <memory>
#include <mutex>

#include "vtp_internal.h"

namespace vtp {

void pbf_step(Ctx &c);

// Hessian-vector product H(r) a of the cubic spline, with gradW = f r:
// H = f I + d r r^T, d = (1/r) df/dr.  In the truncated-power form used by the
// forward passes, |gradW| = G = gt t^2 + gu u^2 with dt/dr = du/dr = -1/h
// where t, u > 0, so dG/dr = -(2/h)(gt t + gu u) and
// d = (dG/dr - G/r) / r^2.
__device__ __forceinline__ void kernel_f_d(float r2, const Cubic &k, float &f, float &d) {
    const float rinv = cubic_rinv(r2);
    const float r = r2 * rinv;
    const float t = unit_clamp(fmaf(-r, k.inv_h, 1.0f)), u = unit_clamp(fmaf(-r, k.inv_h, 0.5f));
    const float G = fmaf(t * t, k.gt, (u * u) * k.gu);
    const float dG = -2.0f * k.inv_h * fmaf(t, k.gt, u * k.gu);
    const bool ok = r2 >= cubic_r2_lo(k);
    f = ok ? G * rinv : 0.0f;
    d = ok ? rinv * rinv * (dG - G * rinv) : 0.0f;
}

__global__ void k_adj_clear_word(uint32_t *w) {
    if (threadIdx.x == 0) *w = 0u;
}

// Terminal adjoints of the device entry point: attach-order float[n][3]
// gradients (either may be null) into the final state's slot order, zero for
// non-fluid particles, as the host entry point fills them.
__global__ void k_adj_terminal(const float *__restrict__ gx, const float *__restrict__ gv,
                               const uint32_t *__restrict__ orig, const int32_t *__restrict__ mat_attach, uint32_t n,
                               float4 *__restrict__ xb, float4 *__restrict__ vb) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t a = orig[i];
    const bool fluid = mat_attach[a] == VTP_MAT_FLUID;
    const size_t o = 3 * (size_t)a;
    xb[i] = (gx && fluid) ? make_float4(gx[o], gx[o + 1], gx[o + 2], 0.0f) : make_float4(0, 0, 0, 0);
    vb[i] = (gv && fluid) ? make_float4(gv[o], gv[o + 1], gv[o + 2], 0.0f) : make_float4(0, 0, 0, 0);
}

// dL/dv0, dL/dx0 (slot order of the rollout's first state) to attach order,
// and dL/dXSPH_k to one float, on the device.
__global__ void k_adj_outputs(const uint32_t *__restrict__ orig, const float4 *__restrict__ xb,
                              const float4 *__restrict__ vb, uint32_t n, float *__restrict__ gv0,
                              float *__restrict__ gx0, const double *__restrict__ kbar, float *__restrict__ gk) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (gk && i == 0) *gk = (float)*kbar;
    if (i >= n) return;
    const size_t o = 3 * (size_t)orig[i];
    if (gv0) { const float4 v = vb[i]; gv0[o] = v.x; gv0[o + 1] = v.y; gv0[o + 2] = v.z; }
    if (gx0) { const float4 x = xb[i]; gx0[o] = x.x; gx0[o + 1] = x.y; gx0[o + 2] = x.z; }
}

__global__ void k_adj_mat_from_orig(const uint32_t *__restrict__ orig, const int32_t *__restrict__ mat_attach,
                                    const uint32_t *__restrict__ scene_attach, uint32_t n, int32_t *__restrict__ mat,
                                    uint32_t *__restrict__ scene) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t a = orig[i];
    mat[i] = mat_attach[a];
    if (scene) scene[i] = scene_attach[a];
}

// u = xbar dt + vbar (adjoint of dv), zero for non-fluid.
// The reverse passes gather a neighbour's data as one 32- or 64-byte record
// (one cache line) instead of one load from each of several arrays; the
// record of slot i is written by the per-particle kernel that produces its
// last field.
//   xvu[4i .. 4i+2] = (x_i, v_i, u_i)           for the XSPH adjoint
//   xw [2i .. 2i+1] = (x_i, w_i)                for lambda_bar
//   rec[4i .. 4i+3] = (x_i, (w_i, l_i), (G_i, rb_i), (sb_i, 0, 0, 0))  for xbar
// Also writes the tick's per-slot materials (mat_attach of the taped slot map),
// which every later reverse pass of the tick reads.
__global__ void k_adj_u(const float4 *__restrict__ xb, const float4 *__restrict__ vb,
                        const uint32_t *__restrict__ orig, const int32_t *__restrict__ mat_attach, uint32_t n,
                        float dt, const float4 *__restrict__ X, const float4 *__restrict__ V,
                        float4 *__restrict__ xvu, int32_t *__restrict__ mat) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t m = mat_attach[orig[i]];
    mat[i] = m;
    float4 a = make_float4(0, 0, 0, 0);
    if (m == VTP_MAT_FLUID) {
        float4 x = xb[i], v = vb[i];
        a = make_float4(x.x * dt + v.x, x.y * dt + v.y, x.z * dt + v.z, 0.0f);
    }
    xvu[4 * (size_t)i] = X[i];
    xvu[4 * (size_t)i + 1] = V[i];
    xvu[4 * (size_t)i + 2] = a;
}

// w = xbar + vbar/dt is the adjoint of dx of a constraint iteration (zero for
// non-fluid).  The kernel that finishes xbar of a slot (XSPH adjoint for the
// last iteration, xbar of iteration m + 1 for iteration m) writes the slot's
// record xw = (x_m, w) in its epilogue.
__device__ __forceinline__ void write_xw(float4 *__restrict__ xw, uint32_t p, const float4 &x, bool fluid,
                                         const float4 &xb, const float4 &vb, float inv_dt) {
    xw[2 * (size_t)p] = x;
    xw[2 * (size_t)p + 1] = fluid ? make_float4(xb.x + vb.x * inv_dt, xb.y + vb.y * inv_dt, xb.z + vb.z * inv_dt, 0.0f)
                                  : make_float4(0, 0, 0, 0);
}

// Transposed XSPH Jacobian:
//   xbar_p += k sum_j [V_j u_p.(v_j - v_p) + V_p u_j.(v_p - v_j)] gradW_pj
//   vbar_p += k V_p sum_j W_pj u_j - k u_p sum_j V_j W_pj
//   kbar   += sum_p u_p . sum_j V_j (v_j - v_p) W_pj
// L lanes per particle (the list is split over them and the sums combined):
// the reverse passes run on small scenes, where one lane per particle leaves
// most of the GPU idle behind each particle's chain of dependent gathers.
template <uint32_t L>
__global__ __launch_bounds__(256) void k_adj_xsph(const float4 *__restrict__ xvu, const int32_t *__restrict__ mat,
                                                  const uint32_t *__restrict__ nbr, const uint32_t *__restrict__ cnt,
                                                  uint32_t n, uint32_t stride, PbfK q, float4 *__restrict__ xb,
                                                  float4 *__restrict__ vb, double *__restrict__ kpart,
                                                  const float4 *__restrict__ Xw, float4 *__restrict__ xw) {
    const uint32_t g = slot_index(), p = g / L, sub = g % L;
    const bool fluid = p < n && mat[p] == VTP_MAT_FLUID;
    if (Xw && p < n && !fluid && sub == 0) write_xw(xw, p, Xw[p], false, make_float4(0, 0, 0, 0),
                                                    make_float4(0, 0, 0, 0), q.inv_dt);
    double kp = 0.0;
    if (fluid) {
        const float4 xp = xvu[4 * (size_t)p], vp = xvu[4 * (size_t)p + 1], up = xvu[4 * (size_t)p + 2];
        const float k = q.xsph_k;
        const uint32_t nn = cnt[p] & ~kDeadBit;
        float ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0, sw = 0, kk = 0;
        struct D { float4 x, v, u; };
        for_each_neighbor_coop<L>(
            nbr, stride, p, nn, sub,
            [&](uint32_t j) { return D{ldg(xvu, 4 * j), ldg(xvu, 4 * j + 1), ldg(xvu, 4 * j + 2)}; },
            [&](const D &e) {
                float dx = xp.x - e.x.x, dy = xp.y - e.x.y, dz = xp.z - e.x.z;
                float W, f;
                cubic_w_grad(dx * dx + dy * dy + dz * dz, q.k, W, f);
                float dvx = e.v.x - vp.x, dvy = e.v.y - vp.y, dvz = e.v.z - vp.z;
                float upd = up.x * dvx + up.y * dvy + up.z * dvz;
                float ujd = -(e.u.x * dvx + e.u.y * dvy + e.u.z * dvz);
                float s = k * (e.v.w * upd + vp.w * ujd) * f;
                ax = fmaf(dx, s, ax);
                ay = fmaf(dy, s, ay);
                az = fmaf(dz, s, az);
                float c = k * vp.w * W;
                bx = fmaf(c, e.u.x, bx);
                by = fmaf(c, e.u.y, by);
                bz = fmaf(c, e.u.z, bz);
                sw = fmaf(e.v.w, W, sw);
                kk = fmaf(e.v.w * W, upd, kk);
            });
        ax = lane_sum<L>(ax);
        ay = lane_sum<L>(ay);
        az = lane_sum<L>(az);
        bx = lane_sum<L>(bx);
        by = lane_sum<L>(by);
        bz = lane_sum<L>(bz);
        sw = lane_sum<L>(sw);
        kk = lane_sum<L>(kk);
        if (sub == 0) {
            float4 a = xb[p], b = vb[p];
            a.x += ax;
            a.y += ay;
            a.z += az;
            b.x += bx - k * sw * up.x;
            b.y += by - k * sw * up.y;
            b.z += bz - k * sw * up.z;
            xb[p] = a;
            vb[p] = b;
            kp = kk;
            if (Xw) write_xw(xw, p, Xw[p], true, a, b, q.inv_dt);
        }
    }
    // one FP64 partial of the XSPH-coefficient adjoint per workgroup, summed
    // in a fixed order by k_adj_kbar (no shared-address atomics: a fixed
    // result, and no serialisation of thousands of atomics on one word)
    __shared__ double s_part[kMaxWavesPerBlock];
    const double ws = wave_sum(kp);
    if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = ws;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (uint32_t w = 0; w < (blockDim.x >> 6); ++w) t += s_part[w];
        kpart[blockIdx.x] = t;
    }
}

// kbar += the partials of one XSPH adjoint launch, summed by 256 strided lanes
// and a fixed tree.
__global__ __launch_bounds__(256) void k_adj_kbar(const double *__restrict__ kpart, uint32_t nparts,
                                                  double *__restrict__ kbar) {
    __shared__ double s_red[256];
    double t = 0.0;
    for (uint32_t k = threadIdx.x; k < nparts; k += 256u) t += kpart[k];
    s_red[threadIdx.x] = t;
    __syncthreads();
    for (uint32_t off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) s_red[threadIdx.x] += s_red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) *kbar += s_red[0];
}

// lambda_bar_i = (1/rho0) sum_j (w_i - w_j) . gradW_ij, then the adjoints of the
// density (rb, zero where clamped) and of S (sb): rs = (rb, sb).
template <uint32_t L>
__global__ __launch_bounds__(256) void k_adj_lambar(const float4 *__restrict__ xw, const float *__restrict__ Cc,
                                                    const float4 *__restrict__ sg, const float *__restrict__ lam,
                                                    const int32_t *__restrict__ mat, const uint32_t *__restrict__ nbr,
                                                    const uint32_t *__restrict__ cnt, uint32_t n, uint32_t stride,
                                                    PbfK q, float4 *__restrict__ rec) {
    const uint32_t g = slot_index(), i = g / L, sub = g % L;
    if (i >= n) return;
    const float4 xi = xw[2 * (size_t)i], wi = xw[2 * (size_t)i + 1], Gi = sg[i];
    float4 *const r = rec + 4 * (size_t)i;
    if (mat[i] != VTP_MAT_FLUID) {
        if (sub == 0) {
            r[0] = xi;
            r[1] = make_float4(wi.x, wi.y, wi.z, lam[i]);
            r[2] = make_float4(Gi.x, Gi.y, Gi.z, 0.0f);
            r[3] = make_float4(0, 0, 0, 0);
        }
        return;
    }
    const uint32_t nn = cnt[i] & ~kDeadBit;
    float lb = 0.0f;
    struct D { float4 x, w; };
    for_each_neighbor_coop<L>(
        nbr, stride, i, nn, sub, [&](uint32_t j) { return D{ldg(xw, 2 * j), ldg(xw, 2 * j + 1)}; },
        [&](const D &e) {
            float dx = xi.x - e.x.x, dy = xi.y - e.x.y, dz = xi.z - e.x.z;
            float f = cubic_grad_factor(dx * dx + dy * dy + dz * dz, q.k);
            lb = fmaf(((wi.x - e.w.x) * dx + (wi.y - e.w.y) * dy + (wi.z - e.w.z) * dz), f, lb);
        });
    lb = lane_sum<L>(lb) * q.inv_rho0;
    if (sub != 0) return;
    r[0] = xi;
    r[1] = make_float4(wi.x, wi.y, wi.z, lam[i]);
    const float C = Cc[i];
    const float den = sg[i].w + 1e-6f;
    const float Cb = -lb / den;
    const float Sb = lb * C / (den * den);
    const float rb = C > 0.0f ? Cb * q.inv_rho0 : 0.0f;
    r[2] = make_float4(Gi.x, Gi.y, Gi.z, rb);
    r[3] = make_float4(Sb, 0, 0, 0);
}

// Transposed constraint-iteration Jacobian:
//   xbar_p += sum_j H_pj [ (l_p + l_j)/rho0 (w_p - w_j)
//                          + 2/rho0^2 (sb_p (gradW_pj + G_p) + sb_j (gradW_pj - G_j)) ]
//            + gradW_pj (rb_p m_j + rb_j m_p)
template <uint32_t L>
__global__ __launch_bounds__(256) void k_adj_xbar(const float4 *__restrict__ rec, const int32_t *__restrict__ mat,
                                                  const uint32_t *__restrict__ nbr, const uint32_t *__restrict__ cnt,
                                                  uint32_t n, uint32_t stride, PbfK q, float4 *__restrict__ xb,
                                                  const float4 *__restrict__ vb, const float4 *__restrict__ Xw,
                                                  float4 *__restrict__ xw) {
    const uint32_t g = slot_index(), p = g / L, sub = g % L;
    if (p >= n) return;
    if (mat[p] != VTP_MAT_FLUID) {
        if (Xw && sub == 0) write_xw(xw, p, Xw[p], false, make_float4(0, 0, 0, 0), make_float4(0, 0, 0, 0), q.inv_dt);
        return;
    }
    const float4 *const own = rec + 4 * (size_t)p;
    const float4 xp = own[0], wl = own[1], gr = own[2];
    const float4 wp = wl, Gp = gr;
    const float2 rp = make_float2(gr.w, own[3].x);
    const float lp = wl.w;
    const float c2 = 2.0f * q.inv_rho0 * q.inv_rho0;
    const uint32_t nn = cnt[p] & ~kDeadBit;
    float ax = 0, ay = 0, az = 0;
    struct D { float4 x, w, G; float2 r; float l; };
    const float *const recf = reinterpret_cast<const float *>(rec);
    for_each_neighbor_coop<L>(
        nbr, stride, p, nn, sub,
        [&](uint32_t j) {
            const float4 a = ldg(rec, 4 * j), b = ldg(rec, 4 * j + 1), g = ldg(rec, 4 * j + 2);
            return D{a, b, g, make_float2(g.w, ldg(recf, 16 * j + 12)), b.w};   // one 64-byte record
        },
        [&](const D &e) {
            float rx = xp.x - e.x.x, ry = xp.y - e.x.y, rz = xp.z - e.x.z;
            float f, d;
            kernel_f_d(rx * rx + ry * ry + rz * rz, q.k, f, d);
            float gx = f * rx, gy = f * ry, gz = f * rz;
            float cl = (lp + e.l) * q.inv_rho0;
            float tx = cl * (wp.x - e.w.x) + c2 * (rp.y * (gx + Gp.x) + e.r.y * (gx - e.G.x));
            float ty = cl * (wp.y - e.w.y) + c2 * (rp.y * (gy + Gp.y) + e.r.y * (gy - e.G.y));
            float tz = cl * (wp.z - e.w.z) + c2 * (rp.y * (gz + Gp.z) + e.r.y * (gz - e.G.z));
            float rt = d * (rx * tx + ry * ty + rz * tz);
            float dm = rp.x * e.x.w + e.r.x * xp.w;
            ax += f * tx + rt * rx + dm * gx;
            ay += f * ty + rt * ry + dm * gy;
            az += f * tz + rt * rz + dm * gz;
        });
    ax = lane_sum<L>(ax);
    ay = lane_sum<L>(ay);
    az = lane_sum<L>(az);
    if (sub != 0) return;
    float4 a = xb[p];
    a.x += ax;
    a.y += ay;
    a.z += az;
    xb[p] = a;
    if (Xw) write_xw(xw, p, Xw[p], true, a, vb[p], q.inv_dt);   // w of the previous iteration
}

// Undo the reorder (x_new[i] = x_old[perm[i]]) and the predict (v1 = v0 + g dt,
// x1 = x0 + v1 dt: vbar0 = vbar1 + dt xbar1, xbar0 = xbar1).
__global__ void k_adj_unpermute_predict(const uint32_t *__restrict__ perm, const float4 *__restrict__ xb,
                                        const float4 *__restrict__ vb, uint32_t n, float dt,
                                        float4 *__restrict__ xo, float4 *__restrict__ vo) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t o = perm[i];
    float4 x = xb[i], v = vb[i];
    xo[o] = x;
    vo[o] = make_float4(v.x + dt * x.x, v.y + dt * x.y, v.z + dt * x.z, 0.0f);
}

// Density loss: L = mean_f (rho_f/rho0 - 1)^2 over the current lists
// (unclamped, self included); rb_i = dL/drho_i.
__global__ void k_dl_rho(const float4 *__restrict__ X, const int32_t *__restrict__ mat,
                         const uint32_t *__restrict__ nbr, const uint32_t *__restrict__ cnt, uint32_t n,
                         uint32_t stride, PbfK q, float inv_nf, float *__restrict__ rb, double *loss) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double e2 = 0.0;
    if (i < n) {
        rb[i] = 0.0f;
        if (mat[i] == VTP_MAT_FLUID) {
            const float4 xi = X[i];
            const uint32_t w = cnt[i];
            const uint32_t nn = w & ~kDeadBit;
            float rho = (w & kDeadBit) ? 0.0f : xi.w * q.k.w0;
            for_each_neighbor(
                nbr, stride, i, nn, [&](uint32_t j) { return ldg(X, j); },
                [&](const float4 &xj) {
                    float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
                    rho = fmaf(xj.w, cubic_w(dx * dx + dy * dy + dz * dz, q.k), rho);
                });
            float e = rho * q.inv_rho0 - 1.0f;
            e2 = (double)e * e;
            rb[i] = 2.0f * e * q.inv_rho0 * inv_nf;
        }
    }
    double s = wave_sum(e2);
    if ((threadIdx.x & 63) == 0 && s != 0.0) atomicAdd(loss, s);
}

// g: slot order (float4), or, when ga is set, ga[orig[p]] in attach order (float[3]).
__global__ void k_dl_grad(const float4 *__restrict__ X, const float *__restrict__ rb, const int32_t *__restrict__ mat,
                          const uint32_t *__restrict__ nbr, const uint32_t *__restrict__ cnt, uint32_t n,
                          uint32_t stride, PbfK q, float4 *__restrict__ g, float *__restrict__ ga,
                          const uint32_t *__restrict__ orig) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    float ax = 0, ay = 0, az = 0;
    if (mat[p] == VTP_MAT_FLUID) {
        const float4 xp = X[p];
        const float rp = rb[p];
        const uint32_t nn = cnt[p] & ~kDeadBit;
        struct D { float4 x; float r; };
        for_each_neighbor(
            nbr, stride, p, nn, [&](uint32_t j) { return D{ldg(X, j), ldg(rb, j)}; },
            [&](const D &e) {
                float dx = xp.x - e.x.x, dy = xp.y - e.x.y, dz = xp.z - e.x.z;
                float f = cubic_grad_factor(dx * dx + dy * dy + dz * dz, q.k) * (rp * e.x.w + e.r * xp.w);
                ax = fmaf(dx, f, ax);
                ay = fmaf(dy, f, ay);
                az = fmaf(dz, f, az);
            });
    }
    if (ga) {
        float *o = ga + 3 * (size_t)orig[p];
        o[0] = ax;
        o[1] = ay;
        o[2] = az;
    } else {
        g[p] = make_float4(ax, ay, az, 0.0f);
    }
}

__global__ void k_dl_clear(double *acc) {
    if (threadIdx.x == 0) *acc = 0.0;
}

// loss = sum / fluid count, in FP64 as the host does it for vtp_pbf_density_loss
__global__ void k_dl_finish(const double *__restrict__ acc, double nf, double *__restrict__ loss) {
    if (threadIdx.x == 0) *loss = nf > 0.0 ? *acc / nf : 0.0;
}

// ------------------------------------------------------------------ host side
struct TapeBufs {
    DevBuf<uint32_t> perm, orig, cnt, nbr;
    DevBuf<float4> x[17];
    DevBuf<float> lam[16], C[16];
    DevBuf<float4> sg[16];
    DevBuf<float4> vlast;
    PbfTapeStep ptrs{};
};

struct Checkpoint {
    DevBuf<float4> x, v;
    DevBuf<uint32_t> orig;
    DevBuf<uint32_t> at_cap;   // neighbour-cap count of the last build (selects the next build's walk)
};

struct AdjointState {
    uint32_t nsteps = 0, every = 0, nseg = 0;
    std::vector<std::unique_ptr<Checkpoint>> ckpt;
    std::vector<std::unique_ptr<TapeBufs>> tape;
    // One captured graph per checkpoint segment (reload, taped re-forward,
    // reverse), replayed by every later backward of the same rollout shape.
    std::vector<hipGraphExec_t> seg_exec;
    hipGraphExec_t tail_exec = nullptr;   // the last segment's reverse alone (its ticks taped by the forward)
    uint64_t seg_epoch = ~0ull;
    uint32_t seg_nsteps = 0, seg_every = 0;
    float4 *canon[4] = {};   // xb, vb, xb2, vb2 as allocated (the reverse ping-pongs them)
    // The forward tapes the ticks of its last segment, so the first segment a
    // backward reverses needs no re-forward.  A backward that re-forwards an
    // earlier segment overwrites those tapes.
    bool tail_taped = false;
    uint32_t capped = 0;   // particles at the cap summed over the forward's builds (0 = lists symmetric)
    void drop_graphs() {
        for (auto &e : seg_exec)
            if (e) (void)hipGraphExecDestroy(e);
        seg_exec.clear();
        if (tail_exec) (void)hipGraphExecDestroy(tail_exec);
        tail_exec = nullptr;
    }
    ~AdjointState() { drop_graphs(); }
    DevBuf<float4> xb, vb, xb2, vb2;
    DevBuf<float4> xvu, xw, rec;   // gather records of the reverse passes
    DevBuf<double> kbar, kpart;   // dL/dXSPH_k and one partial per XSPH-adjoint workgroup
    DevBuf<int32_t> mat;
    bool valid = false;
};

static AdjointState &adj(Ctx &c) {
    if (!c.adj_state) c.adj_state = std::shared_ptr<void>(new AdjointState(), [](void *p) { delete (AdjointState *)p; });
    return *static_cast<AdjointState *>(c.adj_state.get());
}

void launch_mat_from_orig(Ctx &c) {
    hipLaunchKernelGGL(k_adj_mat_from_orig, dim3(div_up(c.n, 256)), dim3(256), 0, c.stream, c.orig.p,
                       c.mat_attach.p, c.scene_attach.p, c.n, c.mat.p, c.scene.p);
}

static void save_ckpt(Ctx &c, Checkpoint &k) {
    const size_t n = c.n;
    VTP_HIP(hipMemcpyAsync(k.x.p, c.X[0].p, n * sizeof(float4), hipMemcpyDeviceToDevice, c.stream));
    VTP_HIP(hipMemcpyAsync(k.v.p, c.V[0].p, n * sizeof(float4), hipMemcpyDeviceToDevice, c.stream));
    VTP_HIP(hipMemcpyAsync(k.orig.p, c.orig.p, n * 4, hipMemcpyDeviceToDevice, c.stream));
    VTP_HIP(hipMemcpyAsync(k.at_cap.p, c.status.p + kStNbrAtCap, 4, hipMemcpyDeviceToDevice, c.stream));
}

static void load_ckpt(Ctx &c, Checkpoint &k) {
    const uint32_t n = c.n;
    VTP_HIP(hipMemcpyAsync(c.X[0].p, k.x.p, n * sizeof(float4), hipMemcpyDeviceToDevice, c.stream));
    VTP_HIP(hipMemcpyAsync(c.V[0].p, k.v.p, n * sizeof(float4), hipMemcpyDeviceToDevice, c.stream));
    VTP_HIP(hipMemcpyAsync(c.orig.p, k.orig.p, (size_t)n * 4, hipMemcpyDeviceToDevice, c.stream));
    VTP_HIP(hipMemcpyAsync(c.status.p + kStNbrAtCap, k.at_cap.p, 4, hipMemcpyDeviceToDevice, c.stream));
    hipLaunchKernelGGL(k_adj_mat_from_orig, dim3(div_up(n, 256)), dim3(256), 0, c.stream, c.orig.p,
                       c.mat_attach.p, c.scene_attach.p, n, c.mat.p, c.scene.p);
}

static void alloc_tape(Ctx &c, TapeBufs &t) {
    const uint32_t n = c.n, K = c.p.iteration_num;
    t.perm.alloc(n);
    t.orig.alloc(n);
    t.cnt.alloc(n);
    t.nbr.alloc(nbr_words(c.ugns.cap, c.ugns.stride));
    for (uint32_t m = 0; m <= K; ++m) t.x[m].alloc(n);
    for (uint32_t m = 0; m < K; ++m) {
        t.lam[m].alloc(n);
        t.C[m].alloc(n);
        t.sg[m].alloc(n);
    }
    t.vlast.alloc(n);
    t.ptrs.perm = t.perm.p;
    t.ptrs.orig = t.orig.p;
    t.ptrs.cnt = t.cnt.p;
    t.ptrs.nbr = t.nbr.p;
    for (uint32_t m = 0; m <= K; ++m) t.ptrs.x[m] = t.x[m].p;
    for (uint32_t m = 0; m < K; ++m) {
        t.ptrs.lam[m] = t.lam[m].p;
        t.ptrs.C[m] = t.C[m].p;
        t.ptrs.sg[m] = t.sg[m].p;
    }
    t.ptrs.v_last = t.vlast.p;
}

// Tapes for one checkpoint segment and the adjoint buffers, allocated once per
// context (and again when the segment grows).
static void ensure_tapes(Ctx &c, AdjointState &A) {
    if (A.tape.size() >= A.every) return;
    const uint32_t n = c.n;
    A.tape.clear();
    for (uint32_t t = 0; t < A.every; ++t) {
        auto tb = std::make_unique<TapeBufs>();
        alloc_tape(c, *tb);
        A.tape.push_back(std::move(tb));
    }
    A.xb.alloc(n);
    A.vb.alloc(n);
    A.xb2.alloc(n);
    A.vb2.alloc(n);
    A.xvu.alloc(4 * (size_t)n);
    A.xw.alloc(2 * (size_t)n);
    A.rec.alloc(4 * (size_t)n);
    A.kbar.alloc(1);
    A.kpart.alloc(div_up(4 * n, 256) + 1);   // workgroups of k_adj_xsph at up to 4 lanes per particle
    A.mat.alloc(n);
    A.canon[0] = A.xb.p;
    A.canon[1] = A.vb.p;
    A.canon[2] = A.xb2.p;
    A.canon[3] = A.vb2.p;
    A.drop_graphs();   // captured against the old tapes
}

// Lanes per particle of the reverse passes: 4 on scenes too small to fill the
// GPU with one lane per particle (measured: 100k particles), else 1.
static uint32_t adj_lanes(const Ctx &c) { return c.n < kSmallScene ? 4u : 1u; }

static void reverse_tick(Ctx &c, AdjointState &A, TapeBufs &t) {
    const uint32_t n = c.n, K = c.p.iteration_num, nb = div_up(n, 256);
    const PbfK q = pbf_constants(c);
    Ugns &u = c.ugns;
    // iteration m's record xw is written by the kernel that finishes xbar
    // before it: the XSPH adjoint for m = K - 1, xbar of m + 1 otherwise
    auto xw_of = [&](int m) -> const float4 * { return m >= 0 ? t.x[m].p : nullptr; };
    c.timer.begin(c.stream, "adj_xsph");
    hipLaunchKernelGGL(k_adj_u, dim3(nb), dim3(256), 0, c.stream, A.xb.p, A.vb.p, t.orig.p, c.mat_attach.p, n, q.dt,
                       t.x[K].p, t.vlast.p, A.xvu.p, A.mat.p);
    const uint32_t L = adj_lanes(c), nbl = div_up(n * L, 256);
    if (L == 4)
        hipLaunchKernelGGL(k_adj_xsph<4>, dim3(nbl), dim3(256), 0, c.stream, A.xvu.p, A.mat.p, t.nbr.p, t.cnt.p, n,
                           u.stride, q, A.xb.p, A.vb.p, A.kpart.p, xw_of((int)K - 1), A.xw.p);
    else
        hipLaunchKernelGGL(k_adj_xsph<1>, dim3(nbl), dim3(256), 0, c.stream, A.xvu.p, A.mat.p, t.nbr.p, t.cnt.p, n,
                           u.stride, q, A.xb.p, A.vb.p, A.kpart.p, xw_of((int)K - 1), A.xw.p);
    hipLaunchKernelGGL(k_adj_kbar, dim3(1), dim3(256), 0, c.stream, A.kpart.p, nbl, A.kbar.p);
    c.timer.end(c.stream);
    for (int m = (int)K - 1; m >= 0; --m) {
        c.timer.begin(c.stream, "adj_iteration");
        if (L == 4) {
            hipLaunchKernelGGL(k_adj_lambar<4>, dim3(nbl), dim3(256), 0, c.stream, A.xw.p, t.C[m].p, t.sg[m].p,
                               t.lam[m].p, A.mat.p, t.nbr.p, t.cnt.p, n, u.stride, q, A.rec.p);
            hipLaunchKernelGGL(k_adj_xbar<4>, dim3(nbl), dim3(256), 0, c.stream, A.rec.p, A.mat.p, t.nbr.p, t.cnt.p,
                               n, u.stride, q, A.xb.p, A.vb.p, xw_of(m - 1), A.xw.p);
        } else {
            hipLaunchKernelGGL(k_adj_lambar<1>, dim3(nbl), dim3(256), 0, c.stream, A.xw.p, t.C[m].p, t.sg[m].p,
                               t.lam[m].p, A.mat.p, t.nbr.p, t.cnt.p, n, u.stride, q, A.rec.p);
            hipLaunchKernelGGL(k_adj_xbar<1>, dim3(nbl), dim3(256), 0, c.stream, A.rec.p, A.mat.p, t.nbr.p, t.cnt.p,
                               n, u.stride, q, A.xb.p, A.vb.p, xw_of(m - 1), A.xw.p);
        }
        c.timer.end(c.stream);
    }
    hipLaunchKernelGGL(k_adj_unpermute_predict, dim3(nb), dim3(256), 0, c.stream, t.perm.p, A.xb.p, A.vb.p, n, q.dt,
                       A.xb2.p, A.vb2.p);
    std::swap(A.xb.p, A.xb2.p);
    std::swap(A.vb.p, A.vb2.p);
}

}  // namespace vtp

using namespace vtp;

template <typename F>
static vtp_status adj_guarded(Ctx &c, F &&f) {
    try {
        c.err.clear();
        return f();
    } catch (const HipError &e) {
        c.err = e.what();
        return e.code == hipErrorOutOfMemory ? VTP_ENOMEM : VTP_EHIP;
    } catch (const std::exception &e) {
        c.err = e.what();
        return VTP_EINVAL;
    }
}

static vtp_status adj_check(Ctx &c) {
    if (c.p.solver != VTP_PBF) { c.err = "rollout gradients exist for PBF only"; return VTP_EINVAL; }
    if (c.p.mass_weighted) { c.err = "the adjoint differentiates the reference (mass-free) formulation only"; return VTP_EINVAL; }
    if (c.p.iteration_num > 16) { c.err = "the adjoint supports at most 16 constraint iterations"; return VTP_EINVAL; }
    // gather records are addressed with 32-bit byte offsets (64 bytes per particle)
    if (c.n > (1u << 26)) { c.err = "the adjoint supports at most 2^26 particles"; return VTP_EINVAL; }
    if (!c.initialized) { c.err = "not initialized"; return VTP_ESTATE; }
    return VTP_OK;
}

extern "C" vtp_status vtp_pbf_rollout_fwd(vtp_ctx *ctx, uint32_t nsteps, uint32_t ckpt_every) {
    if (!ctx) return VTP_EINVAL;
    Ctx &c = ctx->c;
    vtp_status s = adj_check(c);
    if (s != VTP_OK) return s;
    if (nsteps == 0 || ckpt_every == 0) { c.err = "nsteps and ckpt_every must be positive"; return VTP_EINVAL; }
    if (c.crashed) { c.err = "solver is crashed"; return VTP_ECRASH; }
    return adj_guarded(c, [&]() -> vtp_status {
        AdjointState &A = adj(c);
        A.valid = false;
        A.nsteps = nsteps;
        A.every = ckpt_every;
        A.nseg = (nsteps + ckpt_every - 1) / ckpt_every;
        if (A.ckpt.size() != A.nseg + 1) {
            A.ckpt.clear();
            for (uint32_t k = 0; k <= A.nseg; ++k) {
                auto cp = std::make_unique<Checkpoint>();
                cp->x.alloc(c.n);
                cp->v.alloc(c.n);
                cp->orig.alloc(c.n);
                cp->at_cap.alloc(1);
                A.ckpt.push_back(std::move(cp));
            }
        }
        ensure_tapes(c, A);
        A.tail_taped = false;
        A.capped = 0;
        // the builds of the rollout add their cap counts to this word
        hipLaunchKernelGGL(k_adj_clear_word, dim3(1), dim3(64), 0, c.stream, c.status.p + kStCapSeen);
        save_ckpt(c, *A.ckpt[0]);
        const uint32_t tail0 = (A.nseg - 1) * ckpt_every;   // first tick of the last segment
        for (uint32_t t = 0; t < nsteps; ++t) {
            if (t >= tail0) pbf_step_taped(c, A.tape[t - tail0]->ptrs);
            else pbf_step(c);
            c.steps += 1;
            c.sim_time += c.p.time_step;
            if ((t + 1) % ckpt_every == 0 || t + 1 == nsteps) save_ckpt(c, *A.ckpt[(t + ckpt_every) / ckpt_every]);
        }
        {   // the last tick's search wrote into its tape: the context gets its lists back
            const TapeBufs &last = *A.tape[nsteps - 1 - tail0];
            Ugns &u = c.ugns;
            VTP_HIP(hipMemcpyAsync(u.nbr.p, last.nbr.p, u.nbr.bytes(), hipMemcpyDeviceToDevice, c.stream));
            VTP_HIP(hipMemcpyAsync(u.cnt.p, last.cnt.p, u.cnt.bytes(), hipMemcpyDeviceToDevice, c.stream));
            VTP_HIP(hipMemcpyAsync(u.perm.p, last.perm.p, u.perm.bytes(), hipMemcpyDeviceToDevice, c.stream));
            c.lists_stale = false;
        }
        read_status(c);
        if (c.h_status[kStCrash] & ~(uint32_t)VTP_CRASH_OUT_OF_GRID) {
            c.crashed = true;
            c.err = "solver crashed during the rollout";
            return VTP_ECRASH;
        }
        A.capped = c.h_status[kStCapSeen];
        if (A.capped) {
            // a list cut by the cap breaks j in N(i) <=> i in N(j), which the
            // reverse passes rely on: no gradient for this rollout
            c.err = "a neighbour list reached maxNeighborNum during the rollout (" + std::to_string(A.capped) +
                    " particle-builds at the cap): the lists are not symmetric, so the adjoint would be wrong; "
                    "raise maxNeighborNum";
            return VTP_ENBRCAP;
        }
        A.valid = true;
        A.tail_taped = true;
        return VTP_OK;
    });
}

// The backward of both entry points; `device`: the gradients in and out are
// device pointers on the context's GPU, else host arrays.
static vtp_status rollout_bwd(vtp_ctx *ctx, const float *grad_x, const float *grad_v, float *grad_v0,
                              float *grad_x0, float *grad_k, bool device) {
    if (!ctx) return VTP_EINVAL;
    Ctx &c = ctx->c;
    vtp_status s = adj_check(c);
    if (s != VTP_OK) return s;
    AdjointState &A = adj(c);
    if (A.capped) {
        c.err = "the last rollout had neighbour lists at maxNeighborNum: no gradient (raise maxNeighborNum)";
        return VTP_ENBRCAP;
    }
    if (!A.valid) { c.err = "vtp_pbf_rollout_bwd needs a preceding vtp_pbf_rollout_fwd"; return VTP_ESTATE; }
    return adj_guarded(c, [&]() -> vtp_status {
        const uint32_t n = c.n;
        if (A.tape.size() < A.every) A.tail_taped = false;   // (re)allocated: nothing taped
        ensure_tapes(c, A);
        if (A.seg_epoch != c.capture_epoch || A.seg_nsteps != A.nsteps || A.seg_every != A.every ||
            A.seg_exec.size() != A.nseg) {
            A.drop_graphs();
            A.seg_exec.assign(A.nseg, nullptr);
            A.seg_epoch = c.capture_epoch;
            A.seg_nsteps = A.nsteps;
            A.seg_every = A.every;
        }
        A.xb.p = A.canon[0];
        A.vb.p = A.canon[1];
        A.xb2.p = A.canon[2];
        A.vb2.p = A.canon[3];
        // terminal adjoints: attach order -> final slot order (non-fluid zero)
        std::vector<uint32_t> orig;
        std::vector<float4> hx, hv;
        if (device) {
            hipLaunchKernelGGL(k_adj_terminal, dim3(div_up(n, 256)), dim3(256), 0, c.stream, grad_x, grad_v,
                               A.ckpt[A.nseg]->orig.p, c.mat_attach.p, n, A.xb.p, A.vb.p);
        } else {
            orig.resize(n);
            hx.resize(n);
            hv.resize(n);
            VTP_HIP(hipMemcpyAsync(orig.data(), A.ckpt[A.nseg]->orig.p, n * 4, hipMemcpyDeviceToHost, c.stream));
            VTP_HIP(hipStreamSynchronize(c.stream));
            for (uint32_t i = 0; i < n; ++i) {
                uint32_t a = orig[i];
                bool fluid = c.h_mat[a] == VTP_MAT_FLUID;
                hx[i] = (grad_x && fluid) ? make_float4(grad_x[3 * a], grad_x[3 * a + 1], grad_x[3 * a + 2], 0)
                                          : make_float4(0, 0, 0, 0);
                hv[i] = (grad_v && fluid) ? make_float4(grad_v[3 * a], grad_v[3 * a + 1], grad_v[3 * a + 2], 0)
                                          : make_float4(0, 0, 0, 0);
            }
            VTP_HIP(hipMemcpyAsync(A.xb.p, hx.data(), n * sizeof(float4), hipMemcpyHostToDevice, c.stream));
            VTP_HIP(hipMemcpyAsync(A.vb.p, hv.data(), n * sizeof(float4), hipMemcpyHostToDevice, c.stream));
        }
        VTP_HIP(hipMemsetAsync(A.kbar.p, 0, sizeof(double), c.stream));
        const bool graphs = !c.timer.enabled;   // stage timing needs eager launches
        for (int seg = (int)A.nseg - 1; seg >= 0; --seg) {
            const uint32_t s0 = (uint32_t)seg * A.every;
            const uint32_t len = std::min(A.every, A.nsteps - s0);
            const bool taped = A.tail_taped && seg == (int)A.nseg - 1;   // the forward's own tapes
            auto enqueue = [&]() {
                if (!taped) {
                    load_ckpt(c, *A.ckpt[seg]);
                    for (uint32_t t = 0; t < len; ++t) pbf_step_taped(c, A.tape[t]->ptrs);
                }
                for (int t = (int)len - 1; t >= 0; --t) reverse_tick(c, A, *A.tape[t]);
            };
            if (!taped) A.tail_taped = false;   // this segment's re-forward overwrites the tapes
            hipGraphExec_t &exec = taped ? A.tail_exec : A.seg_exec[seg];
            if (!graphs) {
                enqueue();
            } else if (!exec) {
                VTP_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
                hipGraph_t g = nullptr;
                try {
                    enqueue();   // the host-side ping-pong swaps happen here, once per reverse tick
                } catch (...) {
                    (void)hipStreamEndCapture(c.stream, &g);   // leave the stream usable
                    if (g) (void)hipGraphDestroy(g);
                    throw;
                }
                VTP_HIP(hipStreamEndCapture(c.stream, &g));
                hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                VTP_HIP(e);
                VTP_HIP(hipGraphLaunch(exec, c.stream));
            } else {
                VTP_HIP(hipGraphLaunch(exec, c.stream));
                for (uint32_t t = 0; t < len; ++t) {   // the swaps the captured reverse ticks made
                    std::swap(A.xb.p, A.xb2.p);
                    std::swap(A.vb.p, A.vb2.p);
                }
            }
        }
        VTP_HIP(hipGetLastError());
        // adjoints are now in the slot order of the rollout's first state
        if (device) {
            hipLaunchKernelGGL(k_adj_outputs, dim3(div_up(n, 256)), dim3(256), 0, c.stream, A.ckpt[0]->orig.p,
                               A.xb.p, A.vb.p, n, grad_v0, grad_x0, A.kbar.p, grad_k);
            VTP_HIP(hipGetLastError());
            load_ckpt(c, *A.ckpt[A.nseg]);   // leave the context at the rollout's final state
            VTP_HIP(hipStreamSynchronize(c.stream));
            c.lists_stale = true;
            return VTP_OK;
        }
        orig.resize(n);
        hx.resize(n);
        hv.resize(n);
        VTP_HIP(hipMemcpyAsync(orig.data(), A.ckpt[0]->orig.p, n * 4, hipMemcpyDeviceToHost, c.stream));
        VTP_HIP(hipMemcpyAsync(hx.data(), A.xb.p, n * sizeof(float4), hipMemcpyDeviceToHost, c.stream));
        VTP_HIP(hipMemcpyAsync(hv.data(), A.vb.p, n * sizeof(float4), hipMemcpyDeviceToHost, c.stream));
        double kb = 0;
        VTP_HIP(hipMemcpyAsync(&kb, A.kbar.p, sizeof(double), hipMemcpyDeviceToHost, c.stream));
        // leave the context at the rollout's final state
        load_ckpt(c, *A.ckpt[A.nseg]);
        VTP_HIP(hipStreamSynchronize(c.stream));
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t a = orig[i];
            if (grad_v0) { grad_v0[3 * a] = hv[i].x; grad_v0[3 * a + 1] = hv[i].y; grad_v0[3 * a + 2] = hv[i].z; }
            if (grad_x0) { grad_x0[3 * a] = hx[i].x; grad_x0[3 * a + 1] = hx[i].y; grad_x0[3 * a + 2] = hx[i].z; }
        }
        if (grad_k) *grad_k = (float)kb;
        c.lists_stale = true;
        return VTP_OK;
    });
}

extern "C" vtp_status vtp_pbf_rollout_bwd(vtp_ctx *ctx, const float *grad_x, const float *grad_v, float *grad_v0,
                                          float *grad_x0, float *grad_k) {
    return rollout_bwd(ctx, grad_x, grad_v, grad_v0, grad_x0, grad_k, false);
}

extern "C" vtp_status vtp_pbf_rollout_bwd_device(vtp_ctx *ctx, const float *d_grad_x, const float *d_grad_v,
                                                 float *d_grad_v0, float *d_grad_x0, float *d_grad_k) {
    return rollout_bwd(ctx, d_grad_x, d_grad_v, d_grad_v0, d_grad_x0, d_grad_k, true);
}

extern "C" vtp_status vtp_pbf_density_loss(vtp_ctx *ctx, double *loss, float *grad_x) {
    if (!ctx) return VTP_EINVAL;
    Ctx &c = ctx->c;
    if (c.p.solver != VTP_PBF) { c.err = "density loss exists for PBF only"; return VTP_EINVAL; }
    if (!c.initialized) { c.err = "not initialized"; return VTP_ESTATE; }
    if (c.lists_stale) { c.err = "neighbour lists are stale after a rollout backward; tick first"; return VTP_ESTATE; }
    return adj_guarded(c, [&]() -> vtp_status {
        const uint32_t n = c.n, nb = div_up(n, 256);
        const PbfK q = pbf_constants(c);
        uint32_t nf = 0;
        for (int32_t m : c.h_mat) nf += m == VTP_MAT_FLUID;
        DevBuf<float> rb;
        DevBuf<float4> g;
        DevBuf<double> L;
        rb.alloc(n);
        g.alloc(n);
        L.alloc(1);
        VTP_HIP(hipMemsetAsync(L.p, 0, sizeof(double), c.stream));
        Ugns &u = c.ugns;
        hipLaunchKernelGGL(k_dl_rho, dim3(nb), dim3(256), 0, c.stream, c.X[0].p, c.mat.p, u.nbr.p, u.cnt.p, n,
                           u.stride, q, nf ? 1.0f / nf : 0.0f, rb.p, L.p);
        hipLaunchKernelGGL(k_dl_grad, dim3(nb), dim3(256), 0, c.stream, c.X[0].p, rb.p, c.mat.p, u.nbr.p, u.cnt.p, n,
                           u.stride, q, g.p, nullptr, nullptr);
        VTP_HIP(hipGetLastError());
        double lsum = 0;
        VTP_HIP(hipMemcpyAsync(&lsum, L.p, sizeof(double), hipMemcpyDeviceToHost, c.stream));
        std::vector<float4> hg(n);
        std::vector<uint32_t> orig(n);
        if (grad_x) {
            VTP_HIP(hipMemcpyAsync(hg.data(), g.p, n * sizeof(float4), hipMemcpyDeviceToHost, c.stream));
            VTP_HIP(hipMemcpyAsync(orig.data(), c.orig.p, n * 4, hipMemcpyDeviceToHost, c.stream));
        }
        VTP_HIP(hipStreamSynchronize(c.stream));
        if (loss) *loss = nf ? lsum / nf : 0.0;
        if (grad_x)
            for (uint32_t i = 0; i < n; ++i) {
                uint32_t a = orig[i];
                grad_x[3 * a] = hg[i].x;
                grad_x[3 * a + 1] = hg[i].y;
                grad_x[3 * a + 2] = hg[i].z;
            }
        return VTP_OK;
    });
}

// ---- allocation-free density loss -----------------------------------------
// The scratch of vtp_pbf_density_loss (the loss accumulator and dL/drho per
// slot) in a fixed device array, so that this variant calls no allocator.
// Layout: [0, 256) the FP64 accumulator, then float rb[kNoAllocMaxParticles].
// Each device holds its own copy at its own address, looked up once per
// device; host code uses that address, never the array's name.
constexpr uint32_t kNoAllocMaxParticles = 1u << 22;
constexpr size_t kNoAllocScratchBytes = 256 + sizeof(float) * (size_t)kNoAllocMaxParticles;   // 16 MiB + 256 B
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kNoAllocMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kNoAllocMaxDevices] = {};

// The device address of `scratch` on `device` (current), looked up on the first call there.
void *scratch_on(int device) {
    if (device < 0 || device >= kNoAllocMaxDevices) throw std::runtime_error("device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        VTP_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) throw std::runtime_error("no device address for the scratch array");
        g_scratch_ptr[device] = ptr;
    }
    return g_scratch_ptr[device];
}
}  // namespace

extern "C" vtp_status vtp_pbf_density_loss_noalloc(vtp_ctx *ctx, double *d_loss, float *d_grad_x) {
    if (!ctx) return VTP_EINVAL;
    Ctx &c = ctx->c;
    if (c.p.solver != VTP_PBF) { c.err = "density loss exists for PBF only"; return VTP_EINVAL; }
    if (!c.initialized) { c.err = "not initialized"; return VTP_ESTATE; }
    if (c.lists_stale) { c.err = "neighbour lists are stale after a rollout backward; tick first"; return VTP_ESTATE; }
    if (!d_loss) { c.err = "vtp_pbf_density_loss_noalloc: null loss pointer"; return VTP_EINVAL; }
    if (c.n > kNoAllocMaxParticles) {
        c.err = "vtp_pbf_density_loss_noalloc: " + std::to_string(c.n) + " particles exceed its scratch of " +
                std::to_string(kNoAllocMaxParticles) + "; use vtp_pbf_density_loss";
        return VTP_EINVAL;
    }
    return adj_guarded(c, [&]() -> vtp_status {
        int cur = 0;
        VTP_HIP(hipGetDevice(&cur));
        if (cur != c.device) VTP_HIP(hipSetDevice(c.device));
        struct Restore {
            int dev, was;
            ~Restore() { if (dev != was) (void)hipSetDevice(was); }
        } restore{c.device, cur};
        unsigned char *base = static_cast<unsigned char *>(scratch_on(c.device));
        double *acc = reinterpret_cast<double *>(base);
        float *rb = reinterpret_cast<float *>(base + 256);
        const uint32_t n = c.n, nb = div_up(n, 256);
        const PbfK q = pbf_constants(c);
        uint32_t nf = 0;
        for (int32_t m : c.h_mat) nf += m == VTP_MAT_FLUID;
        Ugns &u = c.ugns;
        // the accumulator keeps the last call's sum: cleared by a kernel
        hipLaunchKernelGGL(k_dl_clear, dim3(1), dim3(64), 0, c.stream, acc);
        hipLaunchKernelGGL(k_dl_rho, dim3(nb), dim3(256), 0, c.stream, c.X[0].p, c.mat.p, u.nbr.p, u.cnt.p, n,
                           u.stride, q, nf ? 1.0f / nf : 0.0f, rb, acc);
        hipLaunchKernelGGL(k_dl_finish, dim3(1), dim3(64), 0, c.stream, acc, (double)nf, d_loss);
        if (d_grad_x)
            hipLaunchKernelGGL(k_dl_grad, dim3(nb), dim3(256), 0, c.stream, c.X[0].p, rb, c.mat.p, u.nbr.p,
                               u.cnt.p, n, u.stride, q, nullptr, d_grad_x, c.orig.p);
        VTP_HIP(hipGetLastError());
        VTP_HIP(hipStreamSynchronize(c.stream));
        return VTP_OK;
    });
}

extern "C" vtp_status vtp_pbf_density_loss_device(vtp_ctx *ctx, double *d_loss, float *d_grad_x) {
    if (!ctx) return VTP_EINVAL;
    Ctx &c = ctx->c;
    if (c.p.solver != VTP_PBF) { c.err = "density loss exists for PBF only"; return VTP_EINVAL; }
    if (!c.initialized) { c.err = "not initialized"; return VTP_ESTATE; }
    if (c.lists_stale) { c.err = "neighbour lists are stale after a rollout backward; tick first"; return VTP_ESTATE; }
    if (!d_loss) { c.err = "vtp_pbf_density_loss_device: null loss pointer"; return VTP_EINVAL; }
    return adj_guarded(c, [&]() -> vtp_status {
        const uint32_t n = c.n, nb = div_up(n, 256);
        const PbfK q = pbf_constants(c);
        uint32_t nf = 0;
        for (int32_t m : c.h_mat) nf += m == VTP_MAT_FLUID;
        DevBuf<float> rb;
        DevBuf<double> acc;
        rb.alloc(n);
        acc.alloc(1);
        Ugns &u = c.ugns;
        hipLaunchKernelGGL(k_dl_clear, dim3(1), dim3(64), 0, c.stream, acc.p);
        hipLaunchKernelGGL(k_dl_rho, dim3(nb), dim3(256), 0, c.stream, c.X[0].p, c.mat.p, u.nbr.p, u.cnt.p, n,
                           u.stride, q, nf ? 1.0f / nf : 0.0f, rb.p, acc.p);
        hipLaunchKernelGGL(k_dl_finish, dim3(1), dim3(64), 0, c.stream, acc.p, (double)nf, d_loss);
        if (d_grad_x)
            hipLaunchKernelGGL(k_dl_grad, dim3(nb), dim3(256), 0, c.stream, c.X[0].p, rb.p, c.mat.p, u.nbr.p,
                               u.cnt.p, n, u.stride, q, nullptr, d_grad_x, c.orig.p);
        VTP_HIP(hipGetLastError());
        VTP_HIP(hipStreamSynchronize(c.stream));   // before the scratch is freed
        return VTP_OK;
    });
}

extern "C" vtp_status vtp_pbf_set_xsph(vtp_ctx *ctx, float xsph_k) {
    if (!ctx) return VTP_EINVAL;
    Ctx &c = ctx->c;
    if (c.p.solver != VTP_PBF) { c.err = "XSPH_k is a PBF parameter"; return VTP_EINVAL; }
    c.p.xsph_k = xsph_k;
    pbf_capture_reset(c);   // the captured graph holds the old constant
    return VTP_OK;
}
