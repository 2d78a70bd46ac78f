// imm.hip - IMM mixture and IMM-CT viscoelastic steps
// (IMMCudaImpl.cu, IMMSolver.cpp:400-479, IMMCTCudaImpl.cu, IMMCTSolver.cpp:408-492).
//
// Per-phase arrays are phase-major, a[k*stride + slot], so every phase of a
// wavefront's particles is one coalesced access.  The carried state (phase
// fractions, phase velocities, rest density, colour, conformation tensor) is
// gathered into the new slot order after each sort.  Passes:
//   precompute (cr, alpha, v* = v, IMM-CT Cd) -> div solve -> pressure to phases
//   -> viscosity + normal -> tension + phase velocities
//   -> [IMM-CT] velocity gradient + rheology + conformation update -> polymer stress acc
//   -> density solve -> pressure to phases + update_pos
//   -> phase fluxes (drift + diffusion) with the negative-fraction retry loop
//   -> fraction update / regularise / rest density / mixture velocity / colour
//   -> boundary rest density (max over fluid neighbours).
// Generalisation: P phases (<= 8) and a configurable polymer phase (the
// reference IMM-CT is two-phase with index 1 hard-coded).
#include "vtp_internal.h"

namespace vtp {

struct ImmK {
    Cubic k;
    float gx, gy, gz;
    float gamma, c_sigma, h6_64;
    float cf, cd0;
    float rho[VTP_MAX_PHASES], mu[VTP_MAX_PHASES];
    float col[VTP_MAX_PHASES][3];
    float vis_base, vis_max, inv_relax, beta0, theta;
    int P, pp, ct;
    float V0;
    int mass_weighted;   // extension: m_j in the polymer-stress term
};

static ImmK imm_constants(const Ctx &c) {
    ImmK q{};
    q.k = c.kern;
    q.gx = c.p.gravity[0];
    q.gy = c.p.gravity[1];
    q.gz = c.p.gravity[2];
    q.gamma = c.p.surface_tension_coefficient;
    const float pi = 3.14159265f;
    q.c_sigma = 32.0f / pi / std::pow(c.h, 9.0f);
    q.h6_64 = std::pow(c.h, 6.0f) / 64.0f;
    q.cf = c.p.cf;
    q.cd0 = c.p.cd;
    q.P = (int)c.p.phase_num;
    for (int k = 0; k < q.P; ++k) {
        q.rho[k] = c.p.phase_rest_density[k];
        q.mu[k] = c.p.phase_rest_viscosity[k];
        for (int a = 0; a < 3; ++a) q.col[k][a] = c.p.phase_rest_color[k][a];
    }
    // the reference scales these by 1000 on the device (IMMCTCudaImpl.cu:58-61)
    q.vis_base = 1000.0f * c.p.solution_vis_base;
    q.vis_max = 1000.0f * c.p.solution_vis_max;
    q.inv_relax = c.p.relaxation_time > 0.f ? 1.0f / c.p.relaxation_time : 0.f;
    q.beta0 = c.p.thinning_exp0;
    q.theta = c.p.rheology_threshold;
    q.pp = c.p.polymer_phase;
    q.ct = c.p.solver == VTP_IMMCT ? 1 : 0;
    q.V0 = c.V0;
    q.mass_weighted = c.p.mass_weighted;
    return q;
}

struct Ph {   // phase-major arrays
    float *frac;
    float4 *u;
    float4 *a;
    float4 *drift;
    float *fin, *fout;
    uint32_t stride;
};

__device__ __forceinline__ float4 f4(float x, float y, float z) { return make_float4(x, y, z, 0.0f); }

// mixture velocity from phase velocities (update_vel_from_phase_vel_cuda,
// IMMCudaImpl.cu:396-415): v = sum_k frac_k u_k, v* = v, drift_k = u_k - v.
template <int MP>
__device__ __forceinline__ float4 mix_velocity(const Ph &ph, uint32_t i, int P, float4 vold, float4 *vadv) {
    float vx = 0, vy = 0, vz = 0;
#pragma unroll
    for (int k = 0; k < MP; ++k) {
        if (k >= P) break;
        float a = ph.frac[(size_t)k * ph.stride + i];
        float4 u = ph.u[(size_t)k * ph.stride + i];
        vx = fmaf(u.x, a, vx);
        vy = fmaf(u.y, a, vy);
        vz = fmaf(u.z, a, vz);
    }
#pragma unroll
    for (int k = 0; k < MP; ++k) {
        if (k >= P) break;
        float4 u = ph.u[(size_t)k * ph.stride + i];
        // w: the fraction, read with the drift by the phase-flux pass
        ph.drift[(size_t)k * ph.stride + i] = make_float4(u.x - vx, u.y - vy, u.z - vz,
                                                          ph.frac[(size_t)k * ph.stride + i]);
    }
    float4 v = make_float4(vx, vy, vz, vold.w);
    vadv[i] = v;
    return v;
}

// ------------------------------------------------------------------ init
// init_data_cuda + prepare_imm (IMMCudaImpl.cu:51-118, 684-701): V = V0,
// u_k = v, rest = sum_k frac_k rho_k, boundary rest raised to the fluid
// neighbours' maximum (in a second pass: race-free), mass = rest * V0 for
// every particle (prepare_imm runs before the boundary volume is known),
// then the boundary volume 1/sum W.  IMM-CT: Q = I, Cd = Cd0, vis = base.
template <int MP>
__global__ void k_imm_init_a(float4 *__restrict__ x, float4 *__restrict__ v, float4 *__restrict__ vadv,
                             float *__restrict__ rest, float *__restrict__ cd, float *__restrict__ vis,
                             float *__restrict__ Q, float4 *__restrict__ color, const int32_t *__restrict__ mat,
                             uint32_t n, ImmK q, Ph ph) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 vi = v[i];
    vi.w = q.V0;
    v[i] = vi;
    vadv[i] = vi;
    float r = 0.0f, cx = 0, cy = 0, cz = 0;
#pragma unroll
    for (int k = 0; k < MP; ++k) {
        if (k >= q.P) break;
        float a = ph.frac[(size_t)k * ph.stride + i];
        r += a * q.rho[k];
        ph.u[(size_t)k * ph.stride + i] = f4(vi.x, vi.y, vi.z);
        ph.drift[(size_t)k * ph.stride + i] = f4(0, 0, 0);
        ph.fin[(size_t)k * ph.stride + i] = 0.0f;
        ph.fout[(size_t)k * ph.stride + i] = 0.0f;
        cx += a * q.col[k][0];
        cy += a * q.col[k][1];
        cz += a * q.col[k][2];
    }
    rest[i] = r;
    color[i] = f4(cx, cy, cz);
    cd[i] = q.cd0;
    if (q.ct) {
        vis[i] = q.vis_base;
#pragma unroll
        for (int a = 0; a < 9; ++a) Q[(size_t)a * ph.stride + i] = 0.0f;   // Q - I: Q starts at I
    }
    (void)x;
    (void)mat;
}

// Boundary rest density: max(own, fluid neighbours) and mass = rest * V.
// With `volume_pass` the boundary volume 1/sum_{same material} W is computed
// afterwards (prepare_imm order: mass uses V before the volume update).
__global__ void k_imm_boundary_rest(float4 *__restrict__ x, float4 *__restrict__ v, const float *__restrict__ rest_in,
                                    float *__restrict__ rest, const int32_t *__restrict__ mat,
                                    const uint32_t *__restrict__ nbr, const uint32_t *__restrict__ cnt, uint32_t n,
                                    uint32_t stride, Cubic k, int volume_pass) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t mi = mat[i];
    const uint32_t nn = cnt[i] & ~kDeadBit;
    if (mi == VTP_MAT_FLUID) {
        if (volume_pass) x[i].w = rest_in[i] * v[i].w;
        return;
    }
    float r = rest_in[i];
    for (uint32_t t = 0; t < nn; ++t) {
        uint32_t j = nbr_entry(nbr, stride, i, t);
        if (mat[j] == VTP_MAT_FLUID) r = fmaxf(r, rest_in[j]);
    }
    rest[i] = r;
    float4 vi = v[i];
    x[i].w = r * vi.w;
    if (volume_pass && mi == VTP_MAT_BOUNDARY) {
        const float4 xi = x[i];
        float d = (cnt[i] & kDeadBit) ? 0.0f : k.w0;
        for (uint32_t t = 0; t < nn; ++t) {
            uint32_t j = nbr_entry(nbr, stride, i, t);
            if (mat[j] != mi) continue;
            float4 xj = x[j];
            float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
            d += cubic_w(dx * dx + dy * dy + dz * dz, k);
        }
        vi.w = 1.0f / d;
        v[i] = vi;
    }
}

// The same maximum once per tick, after the phase update (volume_pass 0 of
// the kernel above): a static slot's rest density becomes the maximum of its
// own part, which the phase update wrote, and its fluid neighbours' rest
// densities, and its mass rest * V.  The fluid rest densities come from the
// signed copy rs (> 0 exactly at fluid slots; the phase update refreshed
// them), one 4-byte gather per neighbour in batches.  rest is updated in
// place without a race: the pass reads rest only at its own static slot and
// rs only at fluid slots, and writes rest only at static slots.
__global__ __launch_bounds__(256) void k_imm_boundary_rest_tick(float4 *__restrict__ x,
                                                                const float4 *__restrict__ v,
                                                                const float *__restrict__ rs,
                                                                float *__restrict__ rest,
                                                                const int32_t *__restrict__ mat,
                                                                const uint32_t *__restrict__ nbr,
                                                                const uint32_t *__restrict__ cnt, uint32_t n,
                                                                uint32_t stride) {
    const uint32_t i = slot_index();
    if (i >= n) return;
    if (mat[i] == VTP_MAT_FLUID) return;
    float r = rest[i];
    const uint32_t nn = cnt[i] & ~kDeadBit;
    for_each_neighbor(nbr, stride, i, nn, [&](uint32_t j) { return ldg(rs, j); },
                      [&](float d) {
                          if (d > 0.0f) r = fmaxf(r, d);   // fluid neighbours only
                      });
    rest[i] = r;
    x[i].w = r * v[i].w;
}

// ------------------------------------------------------------------ precompute
// compute_compression_ratio + compute_df_beta (IMMCudaImpl.cu:120-175), v* = v,
// and for IMM-CT update_Cd_cuda (IMMCTCudaImpl.cu:782-804):
// Cd = Cd0 + (1 - Cd0) sum_{j same material, j != i} V_j frac_j,p W_ij.
__global__ __launch_bounds__(256) void k_imm_precompute(const float4 *__restrict__ x, const float4 *__restrict__ v,
                                                        float4 *__restrict__ vadv, const int32_t *__restrict__ mat,
                                                        const uint32_t *__restrict__ nbr,
                                                        const uint32_t *__restrict__ cnt, uint32_t n,
                                                        uint32_t stride, ImmK q, const float4 *__restrict__ aux,
                                                        float *__restrict__ cr, float *__restrict__ alpha,
                                                        float *__restrict__ cd, float4 *__restrict__ g1,
                                                        const uint32_t *__restrict__ obj_of,
                                                        float4 *__restrict__ sconst) {
    uint32_t i = slot_index();
    if (i >= n) return;
    const float4 vi = v[i];
    vadv[i] = vi;
    const int32_t mi = mat[i];
    if (mi != VTP_MAT_FLUID) {
        cr[i] = 0.0f;
        alpha[i] = 0.0f;
        g1[i] = make_float4(0, 0, 0, 0);
        sconst[i] = make_float4(0.0f, 0.0f, vi.w, 0.0f);   // alpha = 0: not fluid
        if (q.ct) cd[i] = q.cd0;
        return;
    }
    const float4 xi = x[i];
    const uint32_t w = cnt[i];
    const uint32_t nn = w & ~kDeadBit;
    float c = (w & kDeadBit) ? 0.0f : vi.w * q.k.w0;
    float a1x = 0, a1y = 0, a1z = 0, a2 = 1e-6f, b = 0.0f;
    // a neighbour: (x_j, m_j) and (V_j, polymer fraction, material)
    struct D { float4 x, a; };
    for_each_neighbor(
        nbr, stride, i, nn, [&](uint32_t j) { return D{ldg(x, j), ldg(aux, j)}; },
        [&](const D &d) {
            float dx = xi.x - d.x.x, dy = xi.y - d.x.y, dz = xi.z - d.x.z;
            float W, f;
            cubic_w_grad(dx * dx + dy * dy + dz * dz, q.k, W, f);
            const float Vj = d.a.x;
            const int32_t mj = __float_as_int(d.a.z);
            c = fmaf(Vj, W, c);
            float gx = dx * f, gy = dy * f, gz = dz * f;
            a1x = fmaf(Vj, gx, a1x);
            a1y = fmaf(Vj, gy, a1y);
            a1z = fmaf(Vj, gz, a1z);
            if (mj == VTP_MAT_FLUID) a2 += (gx * gx + gy * gy + gz * gz) * Vj * Vj / d.x.w;
            if (mj == mi) b = fmaf(Vj * d.a.y, W, b);
        });
    cr[i] = c;
    const float al = fmaxf((a1x * a1x + a1y * a1y + a1z * a1z) / xi.w + a2, 1e-6f);
    alpha[i] = al;
    sconst[i] = make_float4(c, al, vi.w, __uint_as_float(obj_of[i]));
    // sum_j V_j gradW_ij for the v* updates, and V_i / m_i (w != 0: fluid)
    g1[i] = make_float4(a1x, a1y, a1z, vi.w / xi.w);
    if (q.ct) cd[i] = q.cd0 + (1.0f - q.cd0) * b;
}

// ------------------------------------------------------------------ pressure -> phases
// apply_pressure_acc (IMMCudaImpl.cu:770-788): acc = (v* - v)/dt; a_k = acc (Cd +
// (1-Cd) rest/rho_k); u_k += a_k dt; mixture velocity.  With `xo` set it also
// runs update_pos (:471-482) in place.
template <int MP>
__global__ void k_imm_pressure_to_phases(float4 *__restrict__ x, float4 *__restrict__ v,
                                         float4 *__restrict__ vadv, float4 *__restrict__ acc,
                                         const int32_t *__restrict__ mat, const float *__restrict__ rest,
                                         const float *__restrict__ cd, uint32_t n, ImmK q, Ph ph,
                                         const SolverScalars *sc, int update_pos, float4 *__restrict__ xveff) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float dt = sc->dt, inv_dt = 1.0f / dt;
    float4 vi = v[i];
    float4 a = vadv[i];
    float4 ac = f4((a.x - vi.x) * inv_dt, (a.y - vi.y) * inv_dt, (a.z - vi.z) * inv_dt);
    acc[i] = ac;
    if (mat[i] != VTP_MAT_FLUID) {
        if (xveff) {
            const float4 xi = x[i];
            xveff[i] = make_float4(xi.x, xi.y, xi.z, 0.0f);
        }
        return;
    }
    const float ci = cd[i], ri = rest[i];
#pragma unroll
    for (int k = 0; k < MP; ++k) {
        if (k >= q.P) break;
        float s = ci + (1.0f - ci) * (ri / q.rho[k]);
        size_t o = (size_t)k * ph.stride + i;
        float4 u = ph.u[o];
        u.x += ac.x * s * dt;
        u.y += ac.y * s * dt;
        u.z += ac.z * s * dt;
        ph.u[o] = u;
    }
    vi = mix_velocity<MP>(ph, i, q.P, vi, vadv);
    v[i] = vi;
    if (update_pos) {
        float4 xi = x[i];
        xi.x += vi.x * dt;
        xi.y += vi.y * dt;
        xi.z += vi.z * dt;
        x[i] = xi;
        // the phase fluxes' record: V_i > 0, no flag is set before the transfer
        if (xveff) xveff[i] = make_float4(xi.x, xi.y, xi.z, vi.w);
    }
}

// ------------------------------------------------------------------ viscosity + normal
// add_phase_acc_gravity + add_phase_acc_vis (IMMCudaImpl.cu:275-321) and
// compute_surface_normal (:323-346):
//   a_k = g + sum_{j same mat, j != i} 10 V_j ((mu_k(1-Cd)(u_k - v_j) + mu_k Cd v_ij) . x_ij) gradW / |x_ij|^2
//   n_i = h sum_{j same mat} m_j/rest_j gradW.
template <int MP>
__global__ __launch_bounds__(256) void k_imm_vis_normal(const float4 *__restrict__ x, const float4 *__restrict__ v,
                                                        const int32_t *__restrict__ mat,
                                                        const float *__restrict__ rest,
                                                        const float4 *__restrict__ xmr,
                                                        const float *__restrict__ cd,
                                                        const uint32_t *__restrict__ nbr,
                                                        const uint32_t *__restrict__ cnt, uint32_t n,
                                                        uint32_t stride, ImmK q, Ph ph,
                                                        float4 *__restrict__ normal) {
    uint32_t i = slot_index();
    if (i >= n) return;
    const int32_t mi = mat[i];
    if (mi != VTP_MAT_FLUID) {
        // slots move between ticks: clear what a fluid particle left (w = 0: not fluid)
        normal[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    const float4 xi = x[i];
    const float4 vi = v[i];
    const float ci = cd[i];
    const uint32_t nn = cnt[i] & ~kDeadBit;
    float4 uk[MP];
    float ax[MP], ay[MP], az[MP];
#pragma unroll
    for (int k = 0; k < MP; ++k) {
        uk[k] = k < q.P ? ph.u[(size_t)k * ph.stride + i] : f4(0, 0, 0);
        ax[k] = q.gx;
        ay[k] = q.gy;
        az[k] = q.gz;
    }
    float nx = 0, ny = 0, nz = 0;
    // a neighbour: (x_j, m_j / signed rest density), > 0 exactly for the
    // fluid ones, the same material as i; and v_j
    struct D { float4 x, v; };
    for_each_neighbor(
        nbr, stride, i, nn, [&](uint32_t j) { return D{ldg(xmr, j), ldg(v, j)}; },
        [&](const D &d) {
            if (!(d.x.w > 0.0f)) return;
            float dx = xi.x - d.x.x, dy = xi.y - d.x.y, dz = xi.z - d.x.z;
            float r2 = dx * dx + dy * dy + dz * dz;
            float f = cubic_grad_factor(r2, q.k);
            float vijx = (vi.x - d.v.x) * dx + (vi.y - d.v.y) * dy + (vi.z - d.v.z) * dz;
            float base = 10.0f * d.v.w * f / r2;
#pragma unroll
            for (int k = 0; k < MP; ++k) {
                if (k >= q.P) break;
                float ukx = (uk[k].x - d.v.x) * dx + (uk[k].y - d.v.y) * dy + (uk[k].z - d.v.z) * dz;
                float s = (q.mu[k] * (1.0f - ci) * ukx + q.mu[k] * ci * vijx) * base;
                ax[k] = fmaf(dx, s, ax[k]);
                ay[k] = fmaf(dy, s, ay[k]);
                az[k] = fmaf(dz, s, az[k]);
            }
            float sn = q.k.h * d.x.w * f;
            nx = fmaf(dx, sn, nx);
            ny = fmaf(dy, sn, ny);
            nz = fmaf(dz, sn, nz);
        });
#pragma unroll
    for (int k = 0; k < MP; ++k)
        if (k < q.P) ph.a[(size_t)k * ph.stride + i] = f4(ax[k], ay[k], az[k]);
    // w: the rest density (> 0), which the surface-tension pass gathers with the normal
    normal[i] = make_float4(nx, ny, nz, rest[i]);
}

// add_phase_acc_surface_tension (IMMCudaImpl.cu:348-381) + phase_acc_2_phase_vel
// + update_vel_from_phase_vel.
template <int MP>
__global__ __launch_bounds__(256) void k_imm_tension_phase_vel(const float4 *__restrict__ x, float4 *__restrict__ v,
                                                               float4 *__restrict__ vadv,
                                                               const int32_t *__restrict__ mat,
                                                               const float *__restrict__ rest,
                                                               const float4 *__restrict__ normal,
                                                               const uint32_t *__restrict__ nbr,
                                                               const uint32_t *__restrict__ cnt, uint32_t n,
                                                               uint32_t stride, ImmK q, Ph ph,
                                                               const SolverScalars *sc) {
    uint32_t i = slot_index();
    if (i >= n) return;
    const int32_t mi = mat[i];
    if (mi != VTP_MAT_FLUID) return;
    const float dt = sc->dt;
    float sx = 0, sy = 0, sz = 0;
    if (q.gamma != 0.0f) {
        const float4 xi = x[i];
        const float4 ni = normal[i];
        const float ri = rest[i];
        const uint32_t nn = cnt[i] & ~kDeadBit;
        // (x_j, m_j) and (n_j, rest density of a fluid slot, 0 for a static one)
        struct D { float4 x, n; };
        for_each_neighbor(
            nbr, stride, i, nn, [&](uint32_t j) { return D{ldg(x, j), ldg(normal, j)}; },
            [&](const D &d) {
                if (!(d.n.w > 0.0f)) return;
                float dx = xi.x - d.x.x, dy = xi.y - d.x.y, dz = xi.z - d.x.z;
                float r = sqrtf(dx * dx + dy * dy + dz * dz);
                float kk = 2.0f * ri / (ri + d.n.w);
                float c1 = -q.gamma * xi.w * d.x.w * surface_tension_c(r, q.k.h, q.c_sigma, q.h6_64) / r;
                float c2 = -q.gamma * xi.w;
                sx += kk * (c1 * dx + c2 * (ni.x - d.n.x));
                sy += kk * (c1 * dy + c2 * (ni.y - d.n.y));
                sz += kk * (c1 * dz + c2 * (ni.z - d.n.z));
            });
    }
#pragma unroll
    for (int k = 0; k < MP; ++k) {
        if (k >= q.P) break;
        size_t o = (size_t)k * ph.stride + i;
        float4 a = ph.a[o];
        float4 u = ph.u[o];
        u.x += (a.x + sx) * dt;
        u.y += (a.y + sy) * dt;
        u.z += (a.z + sz) * dt;
        ph.u[o] = u;
    }
    v[i] = mix_velocity<MP>(ph, i, q.P, v[i], vadv);
}

// ------------------------------------------------------------------ IMM-CT
// compute_vel_grad (IMMCTCudaImpl.cu:687-714) + compute_rheological_factors
// (:806-821) + update_ct (:716-744), one pass:
//   L = sum_{j fluid} m_j/rest_j (v_j - v_i) (x) gradW_ij   (row from v, column from gradW),
//   beta = beta0 max(0, theta - a_p)/(theta + 1e-6),  kappa = (1 - a_p)^-1.2,
//   Q += dt (Q L + L^T Q - (Q - I)/lambda - beta (Q - I) Q),
//   D = (L + L^T)/2, shear = sqrt(tr(D)^2 / 2),
//   mu = min(base kappa, max);  mu = base + (mu - base)/(1 + (10 shear)^5) min(1, a_p/(theta + 1e-6)),
//   stress = mu (Q - I).
__global__ __launch_bounds__(256) void k_immct_conformation(const float4 *__restrict__ x,
                                                            const float4 *__restrict__ v,
                                                            const int32_t *__restrict__ mat,
                                                            const float *__restrict__ rs,
                                                            const float4 *__restrict__ xmr,
                                                            const uint32_t *__restrict__ nbr,
                                                            const uint32_t *__restrict__ cnt, uint32_t n,
                                                            uint32_t stride, ImmK q, const float *__restrict__ fpoly,
                                                            float *__restrict__ Q, float *__restrict__ Lout,
                                                            float *__restrict__ stress, float *__restrict__ vis,
                                                            float4 *__restrict__ srec,
                                                            const SolverScalars *sc) {
    uint32_t i = slot_index();
    if (i >= n) return;
    if (mat[i] != VTP_MAT_FLUID) {
        vis[i] = q.vis_base;
        srec[3 * (size_t)i] = make_float4(0, 0, 0, 0);
        srec[3 * (size_t)i + 1] = make_float4(0, 0, 0, 0);
        srec[3 * (size_t)i + 2] = make_float4(0, 0, 0, 0);   // y = 0: not fluid
        return;
    }
    const float dt = sc->dt;
    const float4 xi = x[i];
    const float4 vi = v[i];
    const uint32_t nn = cnt[i] & ~kDeadBit;
    float L[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    // a neighbour: (x_j, m_j / signed rest density), > 0 exactly for the
    // fluid ones; and v_j
    struct D { float4 x, v; };
    for_each_neighbor(
        nbr, stride, i, nn, [&](uint32_t j) { return D{ldg(xmr, j), ldg(v, j)}; },
        [&](const D &d) {
            if (!(d.x.w > 0.0f)) return;
            float dx = xi.x - d.x.x, dy = xi.y - d.x.y, dz = xi.z - d.x.z;
            float f = cubic_grad_factor(dx * dx + dy * dy + dz * dz, q.k) * d.x.w;
            float gx = dx * f, gy = dy * f, gz = dz * f;
            float ux = d.v.x - vi.x, uy = d.v.y - vi.y, uz = d.v.z - vi.z;
            L[0] = fmaf(ux, gx, L[0]); L[1] = fmaf(ux, gy, L[1]); L[2] = fmaf(ux, gz, L[2]);
            L[3] = fmaf(uy, gx, L[3]); L[4] = fmaf(uy, gy, L[4]); L[5] = fmaf(uy, gz, L[5]);
            L[6] = fmaf(uz, gx, L[6]); L[7] = fmaf(uz, gy, L[7]); L[8] = fmaf(uz, gz, L[8]);
        });
    // The state holds Q - I, not Q: near equilibrium Q - I is orders of
    // magnitude below 1, and FP32 keeps its full relative precision only
    // when it is stored as such (the stress mu (Q - I) would otherwise carry
    // the rounding of 1 + (Q - I)).
    const size_t st = stride;
    float QI[9], Qm[9];
#pragma unroll
    for (int a = 0; a < 9; ++a) {
        QI[a] = Q[a * st + i];
        Qm[a] = QI[a] + ((a == 0 || a == 4 || a == 8) ? 1.0f : 0.0f);
    }
    const float ap = fpoly[i];
    const float beta = q.beta0 * fmaxf(0.0f, q.theta - ap) / (q.theta + 1e-6f);
    const float kap = powf(1.0f - ap, -1.2f);
    float Qn[9];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
            float ql = 0, ltq = 0, qiq = 0;
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                ql += Qm[3 * r + m] * L[3 * m + cc];      // (Q L)_rc
                ltq += L[3 * m + r] * Qm[3 * m + cc];     // (L^T Q)_rc
                qiq += QI[3 * r + m] * Qm[3 * m + cc];    // ((Q - I) Q)_rc
            }
            float dq = ql + ltq - q.inv_relax * QI[3 * r + cc] - beta * qiq;
            Qn[3 * r + cc] = QI[3 * r + cc] + dq * dt;   // the new Q - I
        }
    const float tr = L[0] + L[4] + L[8];   // trace of D equals trace of L
    const float shear = sqrtf(0.5f * tr * tr);
    float mu = fminf(q.vis_base * kap, q.vis_max);
    float s5 = 10.0f * shear;
    s5 = s5 * s5 * s5 * s5 * s5;
    mu = q.vis_base + (mu - q.vis_base) / (1.0f + s5) * fminf(1.0f, ap / (q.theta + 1e-6f));
    vis[i] = mu;
    float S[9];
#pragma unroll
    for (int a = 0; a < 9; ++a) {
        S[a] = mu * Qn[a];
        Q[a * st + i] = Qn[a];
        Lout[a * st + i] = L[a];
        stress[a * st + i] = S[a];
    }
    // the polymer-stress pass gathers S / rest^2 of a neighbour as one 48-byte record
    const float ri = rs[i];   // i is fluid: its signed rest density is its rest density
    const float ir2 = 1.0f / (ri * ri);
    srec[3 * (size_t)i] = make_float4(S[0] * ir2, S[1] * ir2, S[2] * ir2, S[3] * ir2);
    srec[3 * (size_t)i + 1] = make_float4(S[4] * ir2, S[5] * ir2, S[6] * ir2, S[7] * ir2);
    srec[3 * (size_t)i + 2] = make_float4(S[8] * ir2, 1.0f, 0.0f, 0.0f);   // y = 1: fluid
}

// add_phase_pct_acc (IMMCTCudaImpl.cu:746-780) + phase velocities + mixture velocity:
//   a = a_p rest_i sum_{j fluid} (S_i/rest_i^2 + S_j/rest_j^2) gradW_ij,
//   u_k += a (Cd + (1-Cd) rest_i/rho_k) dt.
template <int MP>
__global__ __launch_bounds__(256) void k_immct_stress_acc(const float4 *__restrict__ x, float4 *__restrict__ v,
                                                          float4 *__restrict__ vadv,
                                                          const int32_t *__restrict__ mat,
                                                          const float *__restrict__ rest,
                                                          const float *__restrict__ cd,
                                                          const float *__restrict__ fpoly,
                                                          const float4 *__restrict__ srec,
                                                          const uint32_t *__restrict__ nbr,
                                                          const uint32_t *__restrict__ cnt, uint32_t n,
                                                          uint32_t stride, ImmK q, Ph ph, const SolverScalars *sc) {
    uint32_t i = slot_index();
    if (i >= n) return;
    if (mat[i] != VTP_MAT_FLUID) return;
    const float dt = sc->dt;
    const float4 xi = x[i];
    const float ri = rest[i];
    float Si[9];
    {
        const float4 a = srec[3 * (size_t)i], b = srec[3 * (size_t)i + 1], c = srec[3 * (size_t)i + 2];
        Si[0] = a.x; Si[1] = a.y; Si[2] = a.z; Si[3] = a.w;
        Si[4] = b.x; Si[5] = b.y; Si[6] = b.z; Si[7] = b.w;
        Si[8] = c.x;
    }
    const uint32_t nn = cnt[i] & ~kDeadBit;
    float ax = 0, ay = 0, az = 0;
    // a neighbour is (x_j, S_j / rest_j^2 with its fluid flag): 64 bytes
    struct D { float4 x, a, b, c; };
    for_each_neighbor(
        nbr, stride, i, nn,
        [&](uint32_t j) { return D{ldg(x, j), ldg(srec, 3 * j), ldg(srec, 3 * j + 1), ldg(srec, 3 * j + 2)}; },
        [&](const D &d) {
            if (d.c.y == 0.0f) return;   // not fluid
            float dx = xi.x - d.x.x, dy = xi.y - d.x.y, dz = xi.z - d.x.z;
            float f = cubic_grad_factor(dx * dx + dy * dy + dz * dz, q.k);
            if (q.mass_weighted) f *= d.x.w;
            float gx = dx * f, gy = dy * f, gz = dz * f;
            const float Sj[9] = {d.a.x, d.a.y, d.a.z, d.a.w, d.b.x, d.b.y, d.b.z, d.b.w, d.c.x};
            float T[9];
#pragma unroll
            for (int a = 0; a < 9; ++a) T[a] = Si[a] + Sj[a];
            ax += T[0] * gx + T[1] * gy + T[2] * gz;
            ay += T[3] * gx + T[4] * gy + T[5] * gz;
            az += T[6] * gx + T[7] * gy + T[8] * gz;
        });
    const float sc_a = fpoly[i] * ri;
    ax *= sc_a;
    ay *= sc_a;
    az *= sc_a;
    const float ci = cd[i];
#pragma unroll
    for (int k = 0; k < MP; ++k) {
        if (k >= q.P) break;
        float s = (ci + (1.0f - ci) * (ri / q.rho[k])) * dt;
        size_t o = (size_t)k * ph.stride + i;
        float4 u = ph.u[o];
        u.x += ax * s;
        u.y += ay * s;
        u.z += az * s;
        ph.u[o] = u;
    }
    v[i] = mix_velocity<MP>(ph, i, q.P, v[i], vadv);
}

// ------------------------------------------------------------------ phase transfer
// update_phase_change_from_drift + ..._from_diffuse (IMMCudaImpl.cu:532-605) into
// fresh in/out accumulators; flagged particles (as i or j) are skipped.
template <int MP>
__global__ __launch_bounds__(256) void k_imm_fluxes(const float4 *__restrict__ xveff,
                                                    const int32_t *__restrict__ mat,
                                                    const uint32_t *__restrict__ flag,
                                                    const uint32_t *__restrict__ nbr,
                                                    const uint32_t *__restrict__ cnt, uint32_t n, uint32_t stride,
                                                    ImmK q, Ph ph, const uint32_t *status, const SolverScalars *sc,
                                                    int retry) {
    if (status[kStDone]) return;
    uint32_t i = slot_index();
    if (i >= n) return;
    const int32_t mi = mat[i];
    if (mi != VTP_MAT_FLUID) return;
    if (retry) {
        // A retry round changes the fluxes of a particle only if it or one of
        // its neighbours was flagged by the check just before (tag = that
        // round's iteration); every other particle keeps the fluxes of the
        // previous round, which came from the same inputs.
        const uint32_t tag = status[kStIter] - 1;
        const uint32_t fi0 = flag[i];
        if (fi0 != tag) {
            if (fi0) return;   // flagged in an earlier round: its zero fluxes stand
            const uint32_t nn = cnt[i] & ~kDeadBit;
            bool dirty = false;
            for_each_neighbor(nbr, stride, i, nn, [&](uint32_t j) { return ldg(flag, j); },
                              [&](uint32_t f) { dirty |= f == tag; });
            if (!dirty) return;
        }
    }
    float fin[MP], fout[MP], fi[MP];
    float4 di[MP];
#pragma unroll
    for (int k = 0; k < MP; ++k) {
        fin[k] = 0.0f;
        fout[k] = 0.0f;
        fi[k] = k < q.P ? ph.frac[(size_t)k * ph.stride + i] : 0.0f;
        di[k] = k < q.P ? ph.drift[(size_t)k * ph.stride + i] : f4(0, 0, 0);
    }
    if (!flag[i]) {
        const float dt = sc->dt;
        const float4 xi = xveff[i];
        const uint32_t nn = cnt[i] & ~kDeadBit;
        // a neighbour: its (x, flux volume) record and per phase (drift,
        // fraction); mix_velocity keeps the fraction in the drift's w.  The
        // flux volume is V for an unflagged fluid slot and 0 for a static or
        // flagged one, the neighbours this pass skips (i is fluid, so same
        // material = fluid).
        struct D { float4 x; float4 pd[MP]; };
        for_each_neighbor(
            nbr, stride, i, nn,
            [&](uint32_t j) {
                D d;
                d.x = ldg(xveff, j);
#pragma unroll
                for (int k = 0; k < MP; ++k)
                    d.pd[k] = k < q.P ? ldg(ph.drift + (size_t)k * ph.stride, j) : make_float4(0, 0, 0, 0);
                return d;
            },
            [&](const D &d) {
                if (d.x.w == 0.0f) return;
                const float Vj = d.x.w;
                float dx = xi.x - d.x.x, dy = xi.y - d.x.y, dz = xi.z - d.x.z;
                float r2 = dx * dx + dy * dy + dz * dz;
                float f = cubic_grad_factor(r2, q.k);
                float gx = dx * f, gy = dy * f, gz = dz * f;
                float factor = f;   // (gradW . x_ij) / |x_ij|^2 = f
#pragma unroll
                for (int k = 0; k < MP; ++k) {
                    if (k >= q.P) break;
                    const float fj = d.pd[k].w;
                    const float4 dj = d.pd[k];
                    float c1 = -dt * Vj * ((fi[k] * di[k].x + fj * dj.x) * gx + (fi[k] * di[k].y + fj * dj.y) * gy +
                                           (fi[k] * di[k].z + fj * dj.z) * gz);
                    // negative terms to fout, the rest to fin, written as
                    // selects of the addend (adding +0 leaves either sum as
                    // it is) so that fin and fout stay in registers
                    fout[k] += c1 < 0.0f ? c1 : 0.0f;
                    fin[k] += c1 < 0.0f ? 0.0f : c1;
                    float c2 = dt * q.cf * (fi[k] - fj) * Vj * factor;
                    fout[k] += c2 < 0.0f ? c2 : 0.0f;
                    fin[k] += c2 < 0.0f ? 0.0f : c2;
                }
            });
    }
#pragma unroll
    for (int k = 0; k < MP; ++k) {
        if (k >= q.P) break;
        ph.fin[(size_t)k * ph.stride + i] = fin[k];
        ph.fout[(size_t)k * ph.stride + i] = fout[k];
    }
}

__global__ void k_phase_begin(uint32_t *status) {
    if (threadIdx.x == 0) {
        status[kStDone] = 0;
        status[kStIter] = 1;
        status[kStNegCount] = 0;
    }
}

// check_negative_cuda (IMMCudaImpl.cu:607-633): flag unflagged fluid particles
// whose fraction would turn negative; count them (counter reset by the
// decide kernel, race-free).
template <int MP>
__global__ void k_imm_check_negative(const int32_t *__restrict__ mat, uint32_t *__restrict__ flag,
                                     float4 *__restrict__ xveff, uint32_t n, ImmK q, Ph ph, uint32_t *status) {
    if (status[kStDone]) return;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool neg = false;
    if (i < n && mat[i] == VTP_MAT_FLUID && !flag[i]) {
#pragma unroll
        for (int k = 0; k < MP; ++k) {
            if (k >= q.P) break;
            size_t o = (size_t)k * ph.stride + i;
            if (ph.frac[o] + ph.fout[o] + ph.fin[o] < 0.0f) neg = true;
        }
        if (neg) {
            flag[i] = status[kStIter];   // nonzero: the round that flagged it
            xveff[i].w = 0.0f;           // and no longer a flux partner
        }
    }
    uint64_t b = __ballot(neg);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(&status[kStNegCount], (uint32_t)__popcll(b));
}

// Loop rule of phase_transfer (:903-932): stop when nothing new was flagged or
// iter > cap; otherwise the fluxes are recomputed and iter increments.
__global__ void k_imm_phase_decide(uint32_t *status, uint32_t cap) {
    if (threadIdx.x != 0 || status[kStDone]) return;
    uint32_t iter = status[kStIter];
    uint32_t neg = status[kStNegCount];
    status[kStNegCount] = 0;
    if (neg == 0 || iter > cap) {
        status[kStDone] = 1;
        status[kStPhaseIter] = iter;
        if (iter > cap) atomicOr(&status[kStCrash], (uint32_t)VTP_CRASH_PHASE_ITER_CAP);
    } else {
        status[kStIter] = iter + 1;
    }
}

// update_phase_change + release_unused_drift_vel + release_negative +
// regularize (x2, phase_transfer and update_mass_and_vel) + rest density and
// mass (fluid part) + update_vel_from_phase_vel + update_color
// (IMMCudaImpl.cu:484-515, 635-676, 934-980).
template <int MP>
__global__ void k_imm_phase_update(float4 *__restrict__ x, float4 *__restrict__ v, float4 *__restrict__ vadv,
                                   const int32_t *__restrict__ mat, uint32_t *__restrict__ flag,
                                   float *__restrict__ rest, float *__restrict__ rs, float4 *__restrict__ color,
                                   uint32_t n, ImmK q, Ph ph) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (mat[i] != VTP_MAT_FLUID) {
        float r = 0.0f;
#pragma unroll
        for (int k = 0; k < MP; ++k)
            if (k < q.P) r += ph.frac[(size_t)k * ph.stride + i] * q.rho[k];
        rest[i] = r;   // own part; the fluid-neighbour maximum follows
        return;
    }
    const float4 vi = v[i];
    const bool fl = flag[i] != 0;
    float a[MP];
    float sum = 0.0f;
#pragma unroll
    for (int k = 0; k < MP; ++k) {
        if (k >= q.P) { a[k] = 0.0f; continue; }
        size_t o = (size_t)k * ph.stride + i;
        a[k] = ph.frac[o] + (ph.fout[o] + ph.fin[o]);
        if (fl) ph.u[o] = f4(vi.x, vi.y, vi.z);
    }
    flag[i] = 0;
#pragma unroll
    for (int k = 0; k < MP; ++k) sum += a[k];
#pragma unroll
    for (int k = 0; k < MP; ++k) a[k] /= sum;
    sum = 0.0f;
#pragma unroll
    for (int k = 0; k < MP; ++k) sum += a[k];
    float r = 0.0f, cx = 0, cy = 0, cz = 0;
#pragma unroll
    for (int k = 0; k < MP; ++k) {
        if (k >= q.P) break;
        a[k] /= sum;
        ph.frac[(size_t)k * ph.stride + i] = a[k];
        r += a[k] * q.rho[k];
        cx += a[k] * q.col[k][0];
        cy += a[k] * q.col[k][1];
        cz += a[k] * q.col[k][2];
    }
    rest[i] = r;
    rs[i] = r;   // fluid: the signed copy is the rest density itself
    x[i].w = r * vi.w;
    v[i] = mix_velocity<MP>(ph, i, q.P, vi, vadv);
    color[i] = f4(cx, cy, cz);
}

// The carried per-slot state into the new slot order, one pass and one read
// of the permutation: phase fractions and phase velocities, the rest density
// and its signed copy rs (+rest for a fluid slot, -rest for a static one,
// rest >= 0), the colour and, for IMM-CT (Q_src non-null), Q - I.  x, v and
// mat are already in the new order.  It also writes the records the
// neighbour passes of the tick gather, constant until the positions move:
//   xmr = (x, m / rs): position, and mass over the signed rest density,
//         > 0 exactly for a fluid slot;
//   aux = (V, polymer fraction, material bits, 0), and the polymer fraction
//         alone in fpoly (IMM-CT: phase pp; 0 otherwise).
template <int MP>
__global__ void k_imm_gather_state(const uint32_t *__restrict__ perm, const int32_t *__restrict__ mat, uint32_t n,
                                   uint32_t stride, uint32_t P, const float *__restrict__ frac_src,
                                   float *__restrict__ frac_dst, const float4 *__restrict__ u_src,
                                   float4 *__restrict__ u_dst, const float *__restrict__ rest_src,
                                   float *__restrict__ rest_dst, float *__restrict__ rs,
                                   const float4 *__restrict__ col_src, float4 *__restrict__ col_dst,
                                   const float *__restrict__ Q_src, float *__restrict__ Q_dst,
                                   const float4 *__restrict__ x, const float4 *__restrict__ v, int pp,
                                   float *__restrict__ fpoly, float4 *__restrict__ xmr, float4 *__restrict__ aux) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t o = perm[i];
    float fp = 0.0f;
#pragma unroll
    for (int k = 0; k < MP; ++k) {
        if (k >= (int)P) break;
        const float f = frac_src[(size_t)k * stride + o];
        frac_dst[(size_t)k * stride + i] = f;
        u_dst[(size_t)k * stride + i] = u_src[(size_t)k * stride + o];
        if (k == pp) fp = f;
    }
    const float r = rest_src[o];
    const int32_t mi = mat[i];
    const float rsi = mi == VTP_MAT_FLUID ? r : -r;
    rest_dst[i] = r;
    rs[i] = rsi;
    col_dst[i] = col_src[o];
    if (Q_src) {
#pragma unroll
        for (int a = 0; a < 9; ++a) Q_dst[(size_t)a * stride + i] = Q_src[(size_t)a * stride + o];
    }
    const float4 xi = x[i];
    xmr[i] = make_float4(xi.x, xi.y, xi.z, xi.w / rsi);
    aux[i] = make_float4(v[i].w, fp, __int_as_float(mi), 0.0f);
    fpoly[i] = fp;
}

// ------------------------------------------------------------------ host side
void common_upload_and_search(Ctx &c);
void common_reorder(Ctx &c, const float4 *xs, float4 *xd, float4 *xd2, const float4 *vs, float4 *vd, float4 *vd2);
void launch_step_begin(Ctx &c);
void launch_hash(Ctx &c, const float4 *x);
void dfsph_allocate_common(Ctx &c);
void df_pressure_solve(Ctx &c, const float4 *x, int which);

static Ph phases(Ctx &c) {
    return Ph{c.frac[0].p, c.uphase[0].p, c.aphase.p, c.drift.p, c.frac_in.p, c.frac_out.p, c.ugns.stride};
}

template <int MP>
static void imm_init_impl(Ctx &c) {
    const uint32_t n = c.n, nb = div_up(n, 256);
    const ImmK q = imm_constants(c);
    Ph ph = phases(c);
    hipLaunchKernelGGL(k_imm_init_a<MP>, dim3(nb), dim3(256), 0, c.stream, c.X[0].p, c.V[0].p, c.vadv.p,
                       c.rest2[0].p, c.cd.p, c.vis_ct.p, c.Q[0].p, c.color4[0].p, c.mat.p, n, q, ph);
    // prepare_imm: rest density (boundary max) and mass with V = V0, then boundary volume
    // the boundary maximum reads a copy of the rest densities (race-free, A3);
    // rest2[1] is the unused half of the rest-density ping-pong here
    VTP_HIP(hipMemcpyAsync(c.rest2[1].p, c.rest2[0].p, n * 4, hipMemcpyDeviceToDevice, c.stream));
    hipLaunchKernelGGL(k_imm_boundary_rest, dim3(nb), dim3(256), 0, c.stream, c.X[0].p, c.V[0].p, c.rest2[1].p,
                       c.rest2[0].p, c.mat.p, c.ugns.nbr.p, c.ugns.cnt.p, n, c.ugns.stride, c.kern, 1);
    VTP_HIP(hipGetLastError());
}

void imm_initialize(Ctx &c) {
    const uint32_t n = c.n;
    const uint32_t P = c.p.phase_num;
    if (P == 0 || P > VTP_MAX_PHASES) throw std::runtime_error("IMM needs 1..8 phases");
    c.P = P;
    dfsph_allocate_common(c);
    common_upload_and_search(c);
    const uint32_t st = c.ugns.stride;
    for (int b = 0; b < 2; ++b) {
        c.frac[b].alloc((size_t)P * st);
        c.uphase[b].alloc((size_t)P * st);
        c.color4[b].alloc(st);
        c.rest2[b].alloc(st);
        if (c.p.solver == VTP_IMMCT) c.Q[b].alloc(9 * (size_t)st);
    }
    c.aphase.alloc((size_t)P * st);
    c.drift.alloc((size_t)P * st);
    c.frac_in.alloc((size_t)P * st);
    c.frac_out.alloc((size_t)P * st);
    c.flag.alloc(st);
    c.xveff.alloc(st);
    c.rs.alloc(st);
    c.xmr.alloc(st);
    c.aux.alloc(st);
    c.cd.alloc(st);
    c.fpoly.alloc(st);
    c.vis_ct.alloc(st);
    c.Lgrad.alloc(9 * (size_t)st);
    c.stress.alloc(9 * (size_t)st);
    c.srec.alloc(3 * (size_t)st);
    if (c.p.solver != VTP_IMMCT) c.Q[0].alloc(9);
    VTP_HIP(hipMemsetAsync(c.flag.p, 0, c.flag.bytes(), c.stream));
    VTP_HIP(hipMemsetAsync(c.Lgrad.p, 0, c.Lgrad.bytes(), c.stream));
    VTP_HIP(hipMemsetAsync(c.stress.p, 0, c.stress.bytes(), c.stream));
    VTP_HIP(hipMemsetAsync(c.aphase.p, 0, c.aphase.bytes(), c.stream));
    // attach-order fractions -> phase-major slot order
    std::vector<uint32_t> orig(n);
    VTP_HIP(hipMemcpyAsync(orig.data(), c.orig.p, n * 4, hipMemcpyDeviceToHost, c.stream));
    VTP_HIP(hipStreamSynchronize(c.stream));
    std::vector<float> fr((size_t)P * st, 0.0f);
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t k = 0; k < P; ++k) fr[(size_t)k * st + i] = c.h_frac[(size_t)orig[i] * P + k];
    VTP_HIP(hipMemcpyAsync(c.frac[0].p, fr.data(), fr.size() * 4, hipMemcpyHostToDevice, c.stream));
    // rest density and colour live in the ping-pong state buffers rest2 / color4
    if (P <= 4) imm_init_impl<4>(c); else imm_init_impl<8>(c);
    VTP_HIP(hipStreamSynchronize(c.stream));
}

template <int MP>
static void imm_step_impl(Ctx &c) {
    const uint32_t n = c.n, nb = div_up(n, 256);
    const ImmK q = imm_constants(c);
    Ugns &u = c.ugns;
    const uint32_t st = u.stride;
    const uint32_t P = c.P;
    launch_step_begin(c);
    // ---- neighbour search; carried state into the new slot order
    c.timer.begin(c.stream, "imm_hash");
    launch_hash(c, c.X[0].p);
    c.timer.end(c.stream);
    u.sort_and_rank(c.orig.p, c.stream, c.timer);
    c.timer.begin(c.stream, "imm_reorder");
    common_reorder(c, c.X[0].p, c.X[1].p, nullptr, c.V[0].p, c.V[1].p, nullptr);
    hipLaunchKernelGGL(k_imm_gather_state<MP>, dim3(nb), dim3(256), 0, c.stream, u.perm.p, c.mat.p, n, st, P,
                       c.frac[0].p, c.frac[1].p, c.uphase[0].p, c.uphase[1].p, c.rest2[0].p, c.rest2[1].p, c.rs.p,
                       c.color4[0].p, c.color4[1].p, q.ct ? c.Q[0].p : nullptr, c.Q[1].p, c.X[1].p, c.V[1].p,
                       q.ct ? q.pp : -1, c.fpoly.p, c.xmr.p, c.aux.p);
    c.timer.end(c.stream);
    std::swap(c.X[0].p, c.X[1].p);
    std::swap(c.V[0].p, c.V[1].p);
    std::swap(c.frac[0].p, c.frac[1].p);
    std::swap(c.uphase[0].p, c.uphase[1].p);
    std::swap(c.rest2[0].p, c.rest2[1].p);
    std::swap(c.color4[0].p, c.color4[1].p);
    if (q.ct) std::swap(c.Q[0].p, c.Q[1].p);
    float4 *x = c.X[0].p, *v = c.V[0].p;
    float *rest = c.rest2[0].p;
    Ph ph = phases(c);
    u.build_neighbors(x, c.stream, c.timer);
    // workgroup size per neighbour pass (the tuner's choice, else "imm_pass")
    auto shape = [&](const char *pass) {
        auto it = c.block_of.find(pass);
        const uint32_t b = it != c.block_of.end() ? it->second : c.block("imm_pass");
        return std::make_pair(dim3(div_up(n, b)), dim3(b));
    };
    const auto s_pre = shape("imm_precompute"), s_vis = shape("imm_vis_normal"),
               s_ten = shape("imm_tension_phase_vel"), s_flux = shape("imm_fluxes");
    c.timer.begin(c.stream, "imm_precompute");
    hipLaunchKernelGGL(k_imm_precompute, s_pre.first, s_pre.second, 0, c.stream, x, v, c.vadv.p, c.mat.p, u.nbr.p, u.cnt.p, n,
                       st, q, c.aux.p, c.cr.p, c.alpha.p, c.cd.p, c.g1.p, c.obj_of.p, c.sconst.p);
    c.timer.end(c.stream);
    df_pressure_solve(c, x, 0);
    c.timer.begin(c.stream, "imm_pressure_to_phases");
    hipLaunchKernelGGL(k_imm_pressure_to_phases<MP>, dim3(nb), dim3(256), 0, c.stream, x, v, c.vadv.p, c.acc.p,
                       c.mat.p, rest, c.cd.p, n, q, ph, c.scal.p, 0, nullptr);
    c.timer.end(c.stream);
    c.timer.begin(c.stream, "imm_vis_normal");
    hipLaunchKernelGGL(k_imm_vis_normal<MP>, s_vis.first, s_vis.second, 0, c.stream, x, v, c.mat.p, rest, c.xmr.p,
                       c.cd.p, u.nbr.p, u.cnt.p, n, st, q, ph, c.normal.p);
    c.timer.end(c.stream);
    c.timer.begin(c.stream, "imm_tension_phase_vel");
    hipLaunchKernelGGL(k_imm_tension_phase_vel<MP>, s_ten.first, s_ten.second, 0, c.stream, x, v, c.vadv.p, c.mat.p, rest,
                       c.normal.p, u.nbr.p, u.cnt.p, n, st, q, ph, c.scal.p);
    c.timer.end(c.stream);
    if (q.ct) {
        c.timer.begin(c.stream, "immct_conformation");
        const auto s_conf = shape("immct_conformation"), s_str = shape("immct_stress_acc");
        hipLaunchKernelGGL(k_immct_conformation, s_conf.first, s_conf.second, 0, c.stream, x, v, c.mat.p, c.rs.p, c.xmr.p,
                           u.nbr.p, u.cnt.p, n, st, q, c.fpoly.p, c.Q[0].p, c.Lgrad.p, c.stress.p, c.vis_ct.p, c.srec.p, c.scal.p);
        c.timer.end(c.stream);
        c.timer.begin(c.stream, "immct_stress_acc");
        hipLaunchKernelGGL(k_immct_stress_acc<MP>, s_str.first, s_str.second, 0, c.stream, x, v, c.vadv.p, c.mat.p, rest,
                           c.cd.p, c.fpoly.p, c.srec.p, u.nbr.p, u.cnt.p, n, st, q, ph, c.scal.p);
        c.timer.end(c.stream);
    }
    df_pressure_solve(c, x, 1);
    c.timer.begin(c.stream, "imm_pressure_update_pos");
    hipLaunchKernelGGL(k_imm_pressure_to_phases<MP>, dim3(nb), dim3(256), 0, c.stream, x, v, c.vadv.p, c.acc.p,
                       c.mat.p, rest, c.cd.p, n, q, ph, c.scal.p, 1, c.xveff.p);
    c.timer.end(c.stream);
    // ---- phase transfer with the negative-fraction retry loop
    const uint32_t cap = c.p.max_solver_iterations;
    hipLaunchKernelGGL(k_phase_begin, dim3(1), dim3(64), 0, c.stream, c.status.p);
    c.timer.begin(c.stream, "imm_fluxes");
    hipLaunchKernelGGL(k_imm_fluxes<MP>, s_flux.first, s_flux.second, 0, c.stream, c.xveff.p, c.mat.p, c.flag.p,
                       u.nbr.p, u.cnt.p, n, st, q, ph, c.status.p, c.scal.p, 0);
    c.timer.end(c.stream);
    // Rounds are enqueued in chunks as in the pressure solves: the previous
    // tick's round count first, then max(1, hint/4) per host check.  Loop
    // iteration k checks round k + 1 and, if a fraction went negative,
    // recomputes the fluxes (round k + 2); with R rounds the checks 0..R-1
    // and the recomputations 0..R-2 did work, and only those stay timed.
    const uint32_t phint = c.iter_hint_phase;
    const uint32_t pstep = c.opt_chunk ? c.opt_chunk : std::max(1u, phint / 4);
    uint32_t chunk = c.opt_chunk ? c.opt_chunk : std::max(1u, phint), launched = 0;
    const size_t tmark = c.timer.mark();
    for (;;) {
        for (uint32_t it = 0; it < chunk && launched <= cap; ++it, ++launched) {
            c.timer.begin(c.stream, "imm_check_negative");
            hipLaunchKernelGGL(k_imm_check_negative<MP>, dim3(nb), dim3(256), 0, c.stream, c.mat.p, c.flag.p,
                               c.xveff.p, n, q, ph, c.status.p);
            hipLaunchKernelGGL(k_imm_phase_decide, dim3(1), dim3(64), 0, c.stream, c.status.p, cap);
            c.timer.end(c.stream);
            c.timer.begin(c.stream, "imm_fluxes");
            hipLaunchKernelGGL(k_imm_fluxes<MP>, s_flux.first, s_flux.second, 0, c.stream, c.xveff.p, c.mat.p,
                               c.flag.p, u.nbr.p, u.cnt.p, n, st, q, ph, c.status.p, c.scal.p,
                               c.opt_full_flux_retry ? 0 : 1);
            c.timer.end(c.stream);
        }
        read_status(c);
        if (c.h_status[kStDone] || launched > cap) break;
        chunk = pstep;
    }
    const uint32_t rounds = std::max(1u, c.h_status[kStPhaseIter]);
    if (c.timer.enabled) c.timer.drop_from(tmark + 2 * (size_t)(rounds - 1) + 1);
    c.iter_hint_phase = rounds;
    c.timer.begin(c.stream, "imm_phase_update");
    hipLaunchKernelGGL(k_imm_phase_update<MP>, dim3(nb), dim3(256), 0, c.stream, x, v, c.vadv.p, c.mat.p, c.flag.p,
                       rest, c.rs.p, c.color4[0].p, n, q, ph);
    hipLaunchKernelGGL(k_imm_boundary_rest_tick, dim3(nb), dim3(256), 0, c.stream, x, v, c.rs.p, rest, c.mat.p,
                       u.nbr.p, u.cnt.p, n, st);
    c.timer.end(c.stream);
    VTP_HIP(hipGetLastError());
}

void imm_step(Ctx &c) {
    if (c.P <= 4) imm_step_impl<4>(c); else imm_step_impl<8>(c);
}

}  // namespace vtp
