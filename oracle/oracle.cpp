// oracle.cpp - FP64 C++/OpenMP restatement of the reference particle-fluid step.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h).  Each function cites the reference
// code it restates (paths relative to the reference root).
//
// Parity pins (what checks this restatement, in tests/):
//   ugns_update      : std::stable_sort + O(N^2) brute-force neighbour sets on
//                      small inputs (test_oracle_ugns.py), exact-lattice
//                      neighbour counts (33 per interior particle).
//   cubic W / gradW  : closed forms, q=0.5 continuity, unit integral by
//                      quadrature, gradient = finite difference of W.
//   PBF              : free fall in closed form (isolated particle), exact
//                      rest-lattice density, two-particle antisymmetry,
//                      momentum conservation of the fluid-only dx pass.
//   DFSPH            : rest-lattice compression ratio closed form, zero
//                      viscosity for uniform velocity, tension antisymmetry,
//                      iteration/residual semantics on hand-built cases.
//   IMM / IMM-CT     : sum of fractions = 1 after regularise, zero diffusion
//                      for uniform fractions, zero drift flux for zero drift,
//                      Q = I fixed point for L = 0; per-phase totals
//                      conserved through ticks, diffusion down the gradient
//                      (test_oracle_phase_transfer.py); ct_update by the simple-
//                      shear steady state, the logistic solution of the
//                      thinning term and the viscosity law's special points
//                      (test_oracle_conformation.py).
//   PBF adjoint      : central finite differences of this same FP64 forward.
//   closed-form lattice sums (test_oracle_pins.py; sums over exact power-of-two
//   lattices written from the reference kernels' formulas, each failing under
//   a planted dropped factor, flipped sign or transposed index):
//                      DFSPH alpha and the first kappa of a compressed block,
//                      kappa and the v* change at a slab face; PBF delta-x of
//                      a slab's top layer and XSPH under a quadratic shear;
//                      the IMM drift flux of a linear drift and the diffusion
//                      flux of a quadratic fraction field; IMM-CT velocity
//                      gradient of a linear field, Cd on a uniform lattice,
//                      the rheological factors at special fractions and the
//                      polymer-stress acceleration at a slab face.
// Every function above has a pin.  Parity unpinned: the reference's nvcc FP32
// contraction choices (no CUDA toolchain here), which this FP64 restatement
// does not model.
// Values the reference leaves undefined (uninitialised boundary lambda and
// kappa_incomp, racy boundary rest density) are defined as zero / race-free.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <map>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct V3 {
    double x = 0, y = 0, z = 0;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(double s, V3 a) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 &operator+=(V3 &a, V3 b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
inline V3 &operator-=(V3 &a, V3 b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; return a; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double length(V3 a) { return std::sqrt(dot(a, a)); }

// 3x3 row-major matrix (Matrix.hpp:95-104 layout).
struct M3 {
    double m[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    static M3 eye() { M3 r; r.m[0] = r.m[4] = r.m[8] = 1; return r; }
    double &operator()(int r, int c) { return m[3 * r + c]; }
    double operator()(int r, int c) const { return m[3 * r + c]; }
};
inline M3 operator+(const M3 &a, const M3 &b) { M3 r; for (int k = 0; k < 9; ++k) r.m[k] = a.m[k] + b.m[k]; return r; }
inline M3 operator-(const M3 &a, const M3 &b) { M3 r; for (int k = 0; k < 9; ++k) r.m[k] = a.m[k] - b.m[k]; return r; }
inline M3 operator*(const M3 &a, double s) { M3 r; for (int k = 0; k < 9; ++k) r.m[k] = a.m[k] * s; return r; }
inline M3 operator*(double s, const M3 &a) { return a * s; }
inline M3 matmul(const M3 &a, const M3 &b) {
    M3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r(i, j) = a(i, 0) * b(0, j) + a(i, 1) * b(1, j) + a(i, 2) * b(2, j);
    return r;
}
inline M3 transpose(const M3 &a) { M3 r; for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r(i, j) = a(j, i); return r; }
inline M3 outer(V3 a, V3 b) {  // Matrix.hpp:482-499: row from a, column from b
    M3 r;
    double av[3] = {a.x, a.y, a.z}, bv[3] = {b.x, b.y, b.z};
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r(i, j) = av[i] * bv[j];
    return r;
}
inline V3 matvec(const M3 &a, V3 v) {  // Matrix.hpp:855-860
    return {a(0, 0) * v.x + a(0, 1) * v.y + a(0, 2) * v.z, a(1, 0) * v.x + a(1, 1) * v.y + a(1, 2) * v.z,
            a(2, 0) * v.x + a(2, 1) * v.y + a(2, 2) * v.z};
}
inline double trace(const M3 &a) { return a(0, 0) + a(1, 1) + a(2, 2); }

// update_ct_cuda (IMMCTCudaImpl.cu:716-744) for one particle: explicit Euler
// step of Q' = Q L + L^T Q - (Q - I)/relax - beta (Q - I) Q, the viscosity
// mu = base + (min(base kappa, max) - base) / (1 + (10 gamma)^5) min(1, a1/(th + 1e-6))
// with gamma = sqrt(tr(D)^2 / 2), D = (L + L^T)/2 (the reference's trace
// squared), and the stress mu (Q - I).  Q and Qout may alias.
inline void ct_update(const M3 &Q, const M3 &L, double dt, double relax, double beta, double kappa, double vis_base,
                      double vis_max, double a1, double th, M3 &Qout, double &mu_out, M3 &stress) {
    const M3 I = M3::eye();
    const M3 dQ = (matmul(Q, L) + matmul(transpose(L), Q) - (1.0 / relax) * (Q - I) - matmul(beta * (Q - I), Q)) * dt;
    const M3 Qn = Q + dQ;
    const M3 D = 0.5 * (L + transpose(L));
    const double shear = std::sqrt(0.5 * trace(D) * trace(D));
    double mu = std::min(vis_base * kappa, vis_max);
    mu = vis_base + (mu - vis_base) / (1 + std::pow(10 * shear, 5)) * std::min(1.0, a1 / (th + 1e-6));
    Qout = Qn;
    mu_out = mu;
    stress = mu * (Qn - I);
}

constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr double kPi = 3.14159265;  // KernelFunctions.cuh:16 literal

// Cubic spline, KernelFunctions.cuh:13-56, in double.
struct Kern {
    double h, sigma;
    explicit Kern(double h_ = 1) : h(h_), sigma(8.0 / kPi / (h_ * h_ * h_)) {}
    double W(V3 r) const {
        double q = length(r) / h;
        if (q > 1) return 0;
        if (q <= 0.5) return sigma * (6 * q * q * q - 6 * q * q + 1);
        double t = 1 - q;
        return sigma * 2 * t * t * t;
    }
    V3 grad(V3 r) const {
        double rn = length(r);
        double q = rn / h;
        if (q < 1e-6 || q > 1) return {};
        V3 gq = r * (1.0 / (rn * h));
        if (q <= 0.5) return gq * (6 * (3 * q * q - 2 * q) * sigma);
        double t = 1 - q;
        return gq * (-6 * t * t * sigma);
    }
    // KernelFunctions.cuh:58-69
    double tension_C(double r) const {
        double c = 32.0 / kPi / std::pow(h, 9);
        if (r * 2 > h && r <= h) return c * std::pow(h - r, 3) * std::pow(r, 3);
        if (r > 0 && r * 2 <= h) return c * (2 * std::pow(h - r, 3) * std::pow(r, 3) - std::pow(h, 6) / 64);
        return 0;
    }
};

struct Sim {
    vtp_params p{};
    uint32_t n = 0, M = 0, S = 1, P = 0;
    // constants (float values of the reference promoted to double)
    float hf = 0, cellf = 0;
    double h = 0, V0 = 0, dt = 0;
    double xsph = 0;   // XSPH coefficient in double (learnable in the adjoint tests)
    Kern k;
    // grid
    int gx = 0, gy = 0, gz = 0;
    uint32_t C = 0;
    // attach-order state
    std::vector<V3> pos, vel;
    std::vector<int32_t> mat;
    std::vector<uint32_t> scene;
    std::vector<uint32_t> obj_start, obj_end;
    std::vector<int32_t> obj_mat;
    // UGNS (reference layout)
    std::vector<uint32_t> key, sorted, cstart, cend, nnum, nbrs;
    std::vector<uint8_t> alive;
    // PBF
    std::vector<double> mass, volume, density, lamb, error;
    std::vector<V3> dxv;
    // DFSPH / IMM
    std::vector<double> rest_density, vis, cr, alpha, delta, kdiv, kinc;
    std::vector<V3> vadv, acc, normal, color;
    // IMM per phase [p*P + k]
    std::vector<double> frac, frac_in, frac_out;
    std::vector<V3> uph, aph, drift;
    std::vector<int> flag;
    std::vector<double> cd, blocking, vis_ct, beta_ct, kappa_ct;
    std::vector<M3> Q, L, stress;
    double vis_base = 0, vis_max = 0;
    // stats
    uint32_t div_it = 0, inc_it = 0, phase_it = 0, crash = 0;
    double div_res = 0, inc_res = 0;
    bool initialized = false;
    // test probe: 1 = a DFSPH tick returns right after its divergence solve
    // (v* of that solve left in vadv, before the pressure is applied)
    int stop_after = 0;
};

template <typename F>
void par_for(uint32_t n, F &&f) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) f((uint32_t)i);
}

// ---------------------------------------------------------------- UGNS
// UGNSCudaImpl.cu:28-81 (cell of a position), evaluated in FP32.
inline int cell_coord(float x, float lb, float cell) {
    volatile float d = x - lb;
    volatile float q = d / cell;
    return (int)std::floor(q);
}

// UGNSCudaImpl.cu:194: 0 < length(pos_i - pos_j) <= cell, FP32,
// dot = (x*x + y*y) + z*z with every operation rounded.
inline bool neighbor_test(const V3 &a, const V3 &b, float cell) {
    volatile float dx = (float)a.x - (float)b.x;
    volatile float dy = (float)a.y - (float)b.y;
    volatile float dz = (float)a.z - (float)b.z;
    volatile float xx = dx * dx, yy = dy * dy, zz = dz * dz;
    volatile float s = xx + yy;
    volatile float d2 = s + zz;
    float len = std::sqrt((float)d2);
    return len > 0.0f && len <= cell;
}

// UniformGridNeighborSearch.cpp:33-60 offset order.
const int kOff[27][3] = {{-1, -1, -1}, {-1, -1, 0}, {-1, -1, 1}, {-1, 0, -1}, {-1, 0, 0}, {-1, 0, 1},
                         {-1, 1, -1},  {-1, 1, 0},  {-1, 1, 1},  {0, -1, -1}, {0, -1, 0}, {0, -1, 1},
                         {0, 0, -1},   {0, 0, 0},   {0, 0, 1},   {0, 1, -1},  {0, 1, 0},  {0, 1, 1},
                         {1, -1, -1},  {1, -1, 0},  {1, -1, 1},  {1, 0, -1},  {1, 0, 0},  {1, 0, 1},
                         {1, 1, -1},   {1, 1, 0},   {1, 1, 1}};

// cu_update, UGNSCudaImpl.cu:205-227: hash -> stable sort -> ranges -> lists.
// Cells with cp == gridSize are rejected (strict bound; the reference's
// inclusive bound at :79-80 aliases rows).  Off-grid particles sort last in
// attach order and get empty lists; they are recomputed every search.
void ugns_update(Sim &s) {
    const uint32_t n = s.n, M = s.M, dead = s.S * s.C;
    s.key.assign(n, dead);
    s.alive.assign(n, 0);
    const float lbx = s.p.sim_space_lb[0], lby = s.p.sim_space_lb[1], lbz = s.p.sim_space_lb[2];
    uint32_t ndead = 0;
    for (uint32_t i = 0; i < n; ++i) {
        float fx = (float)s.pos[i].x, fy = (float)s.pos[i].y, fz = (float)s.pos[i].z;
        if (!std::isfinite(fx) || !std::isfinite(fy) || !std::isfinite(fz)) {
            ++ndead;
            s.crash |= VTP_CRASH_NAN;
            continue;
        }
        int cx = cell_coord(fx, lbx, s.cellf), cy = cell_coord(fy, lby, s.cellf), cz = cell_coord(fz, lbz, s.cellf);
        if (cx < 0 || cy < 0 || cz < 0 || cx >= s.gx || cy >= s.gy || cz >= s.gz) {
            ++ndead;
            s.crash |= VTP_CRASH_OUT_OF_GRID;
            continue;
        }
        s.key[i] = s.scene[i] * s.C + (uint32_t)cz * (uint32_t)(s.gy * s.gx) + (uint32_t)cy * (uint32_t)s.gx +
                   (uint32_t)cx;
        s.alive[i] = 1;
    }
    (void)ndead;
    s.sorted.resize(n);
    std::iota(s.sorted.begin(), s.sorted.end(), 0u);
    std::stable_sort(s.sorted.begin(), s.sorted.end(),
                     [&](uint32_t a, uint32_t b) { return s.key[a] < s.key[b]; });
    s.cstart.assign((size_t)dead + 1, kNone);
    s.cend.assign((size_t)dead + 1, kNone);
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t kk = s.key[s.sorted[i]];
        if (i == 0 || s.key[s.sorted[i - 1]] != kk) s.cstart[kk] = i;
        if (i == n - 1 || s.key[s.sorted[i + 1]] != kk) s.cend[kk] = i;
    }
    s.nnum.assign(n, 0);
    s.nbrs.assign((size_t)n * M, kNone);
    par_for(n, [&](uint32_t i) {
        uint32_t pi = s.sorted[i];
        if (!s.alive[pi]) return;
        uint32_t kk = s.key[pi];
        uint32_t sc = kk / s.C, local = kk % s.C;
        int cz = (int)(local / (uint32_t)(s.gx * s.gy));
        int cy = (int)((local % (uint32_t)(s.gx * s.gy)) / (uint32_t)s.gx);
        int cx = (int)(local % (uint32_t)s.gx);
        uint32_t *row = &s.nbrs[(size_t)pi * M];
        uint32_t cnt = 1;
        row[0] = pi;
        for (int t = 0; t < 27; ++t) {
            int X = cx + kOff[t][0], Y = cy + kOff[t][1], Z = cz + kOff[t][2];
            if (X < 0 || Y < 0 || Z < 0 || X >= s.gx || Y >= s.gy || Z >= s.gz) continue;
            uint32_t cid = sc * s.C + (uint32_t)Z * (uint32_t)(s.gx * s.gy) + (uint32_t)Y * (uint32_t)s.gx + (uint32_t)X;
            if (s.cstart[cid] == kNone) continue;
            for (uint32_t j = s.cstart[cid]; j <= s.cend[cid]; ++j) {
                uint32_t pj = s.sorted[j];
                if (pj == pi || !s.alive[pj]) continue;
                if (neighbor_test(s.pos[pi], s.pos[pj], s.cellf) && cnt < M) row[cnt++] = pj;
            }
        }
        s.nnum[pi] = cnt;
    });
}

template <typename F>
inline void for_nbrs(const Sim &s, uint32_t pi, F &&f) {
    const uint32_t *row = &s.nbrs[(size_t)pi * s.M];
    for (uint32_t t = 0; t < s.M && row[t] != kNone; ++t) f(row[t]);
}

// ---------------------------------------------------------------- PBF
// PBFCudaImpl.cu:34-68 (init_cuda, compute_rigid_particle_volume_cuda).
void pbf_init(Sim &s) {
    const double rho0f = s.p.f_rest_density, rho0b = s.p.b_rest_density;
    s.mass.assign(s.n, rho0f * s.V0);
    s.volume.assign(s.n, s.V0);
    s.density.assign(s.n, 0);
    s.lamb.assign(s.n, 0);   // boundary lambda is never written by the reference: defined 0
    s.error.assign(s.n, 0);
    s.dxv.assign(s.n, V3{});
    par_for(s.n, [&](uint32_t i) {
        if (s.mat[i] != VTP_MAT_BOUNDARY) return;
        double d = 0;
        for_nbrs(s, i, [&](uint32_t j) { if (s.mat[j] == s.mat[i]) d += s.k.W(s.pos[i] - s.pos[j]); });
        s.volume[i] = 1.0 / d;
        s.mass[i] = rho0b * s.volume[i];
    });
}

// PBFSolver.cpp:326-376 with the kernels of PBFCudaImpl.cu:70-194.
void pbf_tick(Sim &s) {
    const double rho0 = s.p.f_rest_density, dt = s.dt;
    const V3 g{s.p.gravity[0], s.p.gravity[1], s.p.gravity[2]};
    par_for(s.n, [&](uint32_t i) {  // apply_ext_force_cuda :146-157
        if (s.mat[i] != VTP_MAT_FLUID) return;
        s.vel[i] += g * dt;
        s.pos[i] += s.vel[i] * dt;
    });
    ugns_update(s);
    for (uint32_t it = 0; it < s.p.iteration_num; ++it) {
        par_for(s.n, [&](uint32_t i) {  // compute_sph_density_and_error_cuda :70-93
            if (s.mat[i] != VTP_MAT_FLUID) return;
            double d = 0;
            for_nbrs(s, i, [&](uint32_t j) { d += s.mass[j] * s.k.W(s.pos[i] - s.pos[j]); });
            if (d < rho0) d = rho0;
            s.density[i] = d;
            s.error[i] = d / rho0 - 1.0;
        });
        par_for(s.n, [&](uint32_t i) {  // update_lamb_cuda :95-122
            if (s.mat[i] != VTP_MAT_FLUID) return;
            V3 eg{};
            double l = 0;
            for_nbrs(s, i, [&](uint32_t j) {
                V3 w = s.k.grad(s.pos[i] - s.pos[j]);
                if (s.p.mass_weighted) w = w * s.mass[j];   // extension: m_j in the gradient
                V3 egj = w * (-1.0 / rho0);
                eg += w;
                l += dot(egj, egj);
            });
            eg = eg * (1.0 / rho0);
            l += dot(eg, eg);
            s.lamb[i] = -s.error[i] / (l + 1e-6);
        });
        par_for(s.n, [&](uint32_t i) {  // compute_dx_cuda :124-144
            if (s.mat[i] != VTP_MAT_FLUID) return;
            V3 d{};
            for_nbrs(s, i, [&](uint32_t j) {
                double mw = s.p.mass_weighted ? s.mass[j] : 1.0;
                d += (mw * (s.lamb[i] + s.lamb[j])) * s.k.grad(s.pos[i] - s.pos[j]);
            });
            s.dxv[i] = d * (1.0 / rho0);
        });
        par_for(s.n, [&](uint32_t i) {  // apply_dx_cuda :159-170
            if (s.mat[i] != VTP_MAT_FLUID) return;
            s.pos[i] += s.dxv[i];
            s.vel[i] += s.dxv[i] * (1.0 / dt);
        });
    }
    par_for(s.n, [&](uint32_t i) {  // XSPH_cuda :172-194
        if (s.mat[i] != VTP_MAT_FLUID) return;
        V3 dv{};
        for_nbrs(s, i, [&](uint32_t j) {
            dv += (s.xsph * s.volume[j] * s.k.W(s.pos[i] - s.pos[j])) * (s.vel[j] - s.vel[i]);
        });
        s.dxv[i] = dv * dt;
    });
    par_for(s.n, [&](uint32_t i) {
        if (s.mat[i] != VTP_MAT_FLUID) return;
        s.pos[i] += s.dxv[i];
        s.vel[i] += s.dxv[i] * (1.0 / dt);
    });
}

// ---------------------------------------------------------------- DFSPH / IMM shared pieces
// Mean over fluid objects of the per-object mean (CommonFunc.cpp:10-21 used at
// DFSPHCudaImpl.cu:472-483), accumulated in double.
double fluid_residual(const Sim &s, const std::vector<double> &v) {
    double total = 0;
    int nf = 0;
    for (size_t o = 0; o < s.obj_start.size(); ++o) {
        if (s.obj_mat[o] != VTP_MAT_FLUID) continue;
        double sum = 0;
        for (uint32_t i = s.obj_start[o]; i < s.obj_end[o]; ++i) sum += v[i];
        uint32_t cnt = s.obj_end[o] - s.obj_start[o];
        total += cnt ? sum / cnt : 0.0;
        ++nf;
    }
    return nf ? total / nf : 0.0;
}

// compute_compression_ratio_cuda + compute_df_beta_cuda (DFSPHCudaImpl.cu:85-140).
// alpha is only consumed for fluid particles; boundary alpha is left 0.
void df_precompute(Sim &s) {
    par_for(s.n, [&](uint32_t i) {
        s.cr[i] = 0;
        s.alpha[i] = 0;
        if (s.mat[i] != VTP_MAT_FLUID) return;
        double c = 0, a2 = 1e-6;
        V3 a1{};
        for_nbrs(s, i, [&](uint32_t j) {
            V3 r = s.pos[i] - s.pos[j];
            c += s.volume[j] * s.k.W(r);
            V3 w = s.k.grad(r);
            a1 += s.volume[j] * w;
            if (s.mat[j] == VTP_MAT_FLUID) a2 += dot(w, w) * s.volume[j] * s.volume[j] / s.mass[j];
        });
        s.cr[i] = c;
        double a = dot(a1, a1) / s.mass[i] + a2;
        s.alpha[i] = a < 1e-6 ? 1e-6 : a;
    });
}

// One pressure solve (vfsph_div / vfsph_incomp, DFSPHCudaImpl.cu:449-503, 556-611):
// do { delta; residual; kappa; v* update } while (res >= thr && iter <= cap).
void df_pressure_solve(Sim &s, double thr, std::vector<double> &kappa, uint32_t &iters, double &res,
                       uint32_t crash_bit) {
    const double dt = s.dt, inv_dt2 = (1.0 / dt) * (1.0 / dt);
    const uint32_t cap = s.p.max_solver_iterations;
    uint32_t iter = 0;
    double r = 0;
    while (true) {
        ++iter;
        par_for(s.n, [&](uint32_t i) {  // compute_delta_compression_ratio + update_..._from_vel_adv :142-180
            if (s.mat[i] != VTP_MAT_FLUID) return;
            double d = s.cr[i] - 1.0;
            for_nbrs(s, i, [&](uint32_t j) {
                d += dot(s.k.grad(s.pos[i] - s.pos[j]), s.vadv[i] - s.vadv[j]) * s.volume[j] * dt;
            });
            s.delta[i] = d < 0 ? 0 : d;
        });
        r = fluid_residual(s, s.delta);
        par_for(s.n, [&](uint32_t i) {  // compute_kappa_*_from_delta_compression_ratio :182-196, 355-367
            kappa[i] = 0;
            if (s.mat[i] != VTP_MAT_FLUID) return;
            kappa[i] = s.delta[i] / s.alpha[i] * inv_dt2 / s.volume[i];
        });
        std::vector<V3> nv(s.vadv);
        par_for(s.n, [&](uint32_t i) {  // vf_update_vel_adv_from_kappa_* :198-220, 369-391
            if (s.mat[i] != VTP_MAT_FLUID) return;
            V3 v = s.vadv[i];
            for_nbrs(s, i, [&](uint32_t j) {
                v -= (dt * s.volume[i] * s.volume[j] / s.mass[i] * (kappa[i] + kappa[j])) *
                     s.k.grad(s.pos[i] - s.pos[j]);
            });
            nv[i] = v;
        });
        s.vadv.swap(nv);
        if (r < thr || iter > cap) break;
    }
    if (iter > cap) s.crash |= crash_bit;
    iters = iter;
    res = r;
}

// ---------------------------------------------------------------- DFSPH
// init_cuda, update_mass_cuda, compute_rigid_volume_cuda (DFSPHCudaImpl.cu:38-83).
void df_init(Sim &s) {
    const uint32_t n = s.n;
    s.volume.assign(n, s.V0);
    s.mass.assign(n, 0);   // boundary mass is never set by the reference: defined 0
    s.kdiv.assign(n, 0);
    s.kinc.assign(n, 0);   // boundary kappa_incomp is never written: defined 0
    s.acc.assign(n, V3{});
    s.vadv = s.vel;
    s.vis.assign(n, s.p.f_rest_viscosity);
    s.rest_density.assign(n, s.p.f_rest_density);
    s.cr.assign(n, 0);
    s.alpha.assign(n, 0);
    s.delta.assign(n, 0);
    s.normal.assign(n, V3{});
    par_for(n, [&](uint32_t i) {
        if (s.mat[i] == VTP_MAT_FLUID) s.mass[i] = s.rest_density[i] * s.volume[i];
    });
    par_for(n, [&](uint32_t i) {
        if (s.mat[i] != VTP_MAT_BOUNDARY) return;
        double d = 0;
        for_nbrs(s, i, [&](uint32_t j) { if (s.mat[j] == s.mat[i]) d += s.k.W(s.pos[i] - s.pos[j]); });
        s.volume[i] = 1.0 / d;
        s.rest_density[i] = s.volume[i] * s.p.b_rest_density;   // the reference's mass-like quirk
    });
}

// apply_pressure_acc (DFSPHCudaImpl.cu:505-521).
void df_apply_pressure(Sim &s) {
    const double dt = s.dt, inv_dt = 1.0 / dt;
    par_for(s.n, [&](uint32_t i) {
        s.acc[i] = V3{};
        if (s.mat[i] != VTP_MAT_FLUID) return;
        s.acc[i] = (s.vadv[i] - s.vel[i]) * inv_dt;
        s.vel[i] += s.acc[i] * dt;
        s.vadv[i] = s.vel[i];
    });
}

// Surface normal (DFSPHCudaImpl.cu:298-321 / IMMCudaImpl.cu:323-346).
void surface_normal(Sim &s) {
    par_for(s.n, [&](uint32_t i) {
        if (s.mat[i] != VTP_MAT_FLUID) return;
        V3 nn{};
        for_nbrs(s, i, [&](uint32_t j) {
            if (s.mat[j] != s.mat[i]) return;
            nn += (s.h * s.mass[j] / s.rest_density[j]) * s.k.grad(s.pos[i] - s.pos[j]);
        });
        s.normal[i] = nn;
    });
}

// Surface tension acceleration (DFSPHCudaImpl.cu:323-353).
V3 tension_acc(const Sim &s, uint32_t i) {
    V3 a{};
    const double gamma = s.p.surface_tension_coefficient;
    for_nbrs(s, i, [&](uint32_t j) {
        if (s.mat[j] != s.mat[i] || j == i) return;
        V3 r = s.pos[i] - s.pos[j];
        double kk = 2 * s.rest_density[i] / (s.rest_density[i] + s.rest_density[j]);
        V3 a1 = (-gamma * s.mass[i] * s.mass[j] * s.k.tension_C(length(r)) / length(r)) * r;
        V3 a2 = (-gamma * s.mass[i]) * (s.normal[i] - s.normal[j]);
        a += kk * (a1 + a2);
    });
    return a;
}

// DFSPHSolver.cpp:339-406.
void df_tick(Sim &s) {
    const double dt = s.dt;
    const V3 g{s.p.gravity[0], s.p.gravity[1], s.p.gravity[2]};
    ugns_update(s);
    df_precompute(s);
    df_pressure_solve(s, s.p.div_free_threshold, s.kdiv, s.div_it, s.div_res, VTP_CRASH_DIV_ITER_CAP);
    if (s.stop_after == 1) return;
    df_apply_pressure(s);
    // dfsph_gravity_vis_surface :523-554
    const double h2_001 = 0.001 * s.h * s.h;
    par_for(s.n, [&](uint32_t i) {
        s.acc[i] = V3{};
        if (s.mat[i] != VTP_MAT_FLUID) return;
        V3 a = g;
        for_nbrs(s, i, [&](uint32_t j) {  // add_acc_explicit_lap_vis_cuda :264-296
            if (s.mat[j] != s.mat[i]) return;
            V3 xij = s.pos[i] - s.pos[j], vij = s.vel[i] - s.vel[j];
            double vis = (s.vis[i] + s.vis[j]) / 2;
            double pi = vis * s.mass[j] / s.rest_density[j] * dot(vij, xij) / (dot(xij, xij) + h2_001);
            a += (10 * pi) * s.k.grad(xij);
        });
        s.acc[i] = a;
    });
    surface_normal(s);
    par_for(s.n, [&](uint32_t i) {
        if (s.mat[i] != VTP_MAT_FLUID) return;
        s.acc[i] += tension_acc(s, i);
    });
    par_for(s.n, [&](uint32_t i) {  // acc_2_vel :241-251
        if (s.mat[i] != VTP_MAT_FLUID) return;
        s.vel[i] += s.acc[i] * dt;
        s.vadv[i] = s.vel[i];
    });
    df_pressure_solve(s, s.p.incomp_threshold, s.kinc, s.inc_it, s.inc_res, VTP_CRASH_INCOMP_ITER_CAP);
    df_apply_pressure(s);
    par_for(s.n, [&](uint32_t i) {  // update_pos_cuda :393-404
        if (s.mat[i] != VTP_MAT_FLUID) return;
        s.pos[i] += s.vel[i] * dt;
    });
}

void imm_init(Sim &s);
void imm_tick(Sim &s);

}  // namespace

// IMM / IMM-CT restatement lives in oracle_imm.inc (same translation unit).
#include "oracle_imm.inc"

struct orc_sim {
    Sim s;
};

static int g_threads = 0;

extern "C" void orc_set_threads(int n) {
    g_threads = n;
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#endif
}

extern "C" int orc_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

extern "C" orc_sim *orc_create(const vtp_params *p, uint32_t n, const float *pos, const float *vel,
                               const float *vol_frac, const uint32_t *scene, uint32_t nobj,
                               const uint32_t *obj_start, const uint32_t *obj_end, const int32_t *obj_mat) {
    if (!p || !pos || n == 0) return nullptr;
    orc_sim *o = new orc_sim();
    Sim &s = o->s;
    s.p = *p;
    if (s.p.max_solver_iterations == 0) s.p.max_solver_iterations = 100;
    s.n = n;
    s.M = p->max_neighbor_num;
    s.P = (p->solver == VTP_IMM || p->solver == VTP_IMMCT) ? p->phase_num : 0;
    s.hf = p->particle_radius * 4.0f;
    s.cellf = s.hf;
    s.h = s.hf;
    s.V0 = std::pow(2.0f * p->particle_radius, 3.0f);
    s.dt = p->time_step;
    s.xsph = p->xsph_k;
    s.k = Kern(s.h);
    s.gx = (int)(uint32_t)std::ceil(p->sim_space_size[0] / s.cellf);
    s.gy = (int)(uint32_t)std::ceil(p->sim_space_size[1] / s.cellf);
    s.gz = (int)(uint32_t)std::ceil(p->sim_space_size[2] / s.cellf);
    s.C = (uint32_t)(s.gx * s.gy * s.gz);
    s.pos.resize(n);
    s.vel.resize(n);
    s.mat.resize(n);
    s.scene.assign(n, 0);
    for (uint32_t i = 0; i < n; ++i) {
        s.pos[i] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
        if (vel) s.vel[i] = {vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]};
        if (scene) s.scene[i] = scene[i];
    }
    s.S = 1;
    for (uint32_t i = 0; i < n; ++i) s.S = std::max(s.S, s.scene[i] + 1);
    for (uint32_t k = 0; k < nobj; ++k) {
        s.obj_start.push_back(obj_start[k]);
        s.obj_end.push_back(obj_end[k]);
        s.obj_mat.push_back(obj_mat[k]);
        for (uint32_t i = obj_start[k]; i < obj_end[k] && i < n; ++i) s.mat[i] = obj_mat[k];
    }
    if (s.P) {
        s.frac.assign((size_t)n * s.P, 0.0);
        if (vol_frac)
            for (size_t i = 0; i < (size_t)n * s.P; ++i) s.frac[i] = vol_frac[i];
    }
    return o;
}

extern "C" int orc_initialize(orc_sim *o) {
    if (!o) return -1;
    Sim &s = o->s;
    ugns_update(s);
    switch (s.p.solver) {
        case VTP_PBF: pbf_init(s); break;
        case VTP_DFSPH: df_init(s); break;
        default: imm_init(s); break;
    }
    s.initialized = true;
    return 0;
}

extern "C" int orc_step(orc_sim *o, uint32_t nsteps) {
    if (!o) return -1;
    Sim &s = o->s;
    if (!s.initialized) orc_initialize(o);
    for (uint32_t t = 0; t < nsteps; ++t) {
        switch (s.p.solver) {
            case VTP_PBF: pbf_tick(s); break;
            case VTP_DFSPH: df_tick(s); break;
            default: imm_tick(s); break;
        }
    }
    return (s.crash & ~(uint32_t)VTP_CRASH_OUT_OF_GRID) ? 1 : 0;
}

static void put3(double *out, const std::vector<V3> &v, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) {
        out[3 * i] = v[i].x;
        out[3 * i + 1] = v[i].y;
        out[3 * i + 2] = v[i].z;
    }
}
static int put1(double *out, const std::vector<double> &v, uint32_t n) {
    if (v.size() < n) return -2;
    std::copy(v.begin(), v.begin() + n, out);
    return 0;
}

extern "C" int orc_read(orc_sim *o, int32_t field, double *out) {
    if (!o || !out) return -1;
    Sim &s = o->s;
    const uint32_t n = s.n;
    switch (field) {
        case VTP_FIELD_POS: put3(out, s.pos, n); return 0;
        case VTP_FIELD_VEL: put3(out, s.vel, n); return 0;
        case VTP_FIELD_MAT: for (uint32_t i = 0; i < n; ++i) out[i] = s.mat[i]; return 0;
        case VTP_FIELD_SCENE: for (uint32_t i = 0; i < n; ++i) out[i] = s.scene[i]; return 0;
        case VTP_FIELD_MASS: return put1(out, s.mass, n);
        case VTP_FIELD_VOLUME: return put1(out, s.volume, n);
        case VTP_FIELD_DENSITY: return put1(out, s.density, n);
        case VTP_FIELD_LAMBDA: return put1(out, s.lamb, n);
        case VTP_FIELD_ERROR: return put1(out, s.error, n);
        case VTP_FIELD_REST_DENSITY: return put1(out, s.rest_density, n);
        case VTP_FIELD_COMPRESSION_RATIO: return put1(out, s.cr, n);
        case VTP_FIELD_DF_ALPHA: return put1(out, s.alpha, n);
        case VTP_FIELD_DELTA_CR: return put1(out, s.delta, n);
        case VTP_FIELD_KAPPA_DIV: return put1(out, s.kdiv, n);
        case VTP_FIELD_KAPPA_INCOMP: return put1(out, s.kinc, n);
        case VTP_FIELD_VEL_ADV: if (s.vadv.size() < n) return -2; put3(out, s.vadv, n); return 0;
        case VTP_FIELD_ACC: if (s.acc.size() < n) return -2; put3(out, s.acc, n); return 0;
        case VTP_FIELD_SURFACE_NORMAL: if (s.normal.size() < n) return -2; put3(out, s.normal, n); return 0;
        case VTP_FIELD_VIS: return put1(out, s.p.solver == VTP_IMMCT ? s.vis_ct : s.vis, n);
        case VTP_FIELD_VOL_FRAC: if (s.frac.size() < (size_t)n * s.P) return -2;
            std::copy(s.frac.begin(), s.frac.begin() + (size_t)n * s.P, out); return 0;
        case VTP_FIELD_VEL_PHASE: if (s.uph.size() < (size_t)n * s.P) return -2; put3(out, s.uph, n * s.P); return 0;
        case VTP_FIELD_VEL_DRIFT: if (s.drift.size() < (size_t)n * s.P) return -2; put3(out, s.drift, n * s.P); return 0;
        case VTP_FIELD_COLOR: if (s.color.size() < n) return -2; put3(out, s.color, n); return 0;
        case VTP_FIELD_CD: return put1(out, s.cd, n);
        case VTP_FIELD_Q: case VTP_FIELD_VEL_GRAD: case VTP_FIELD_VISE_STRESS: {
            const std::vector<M3> &m = field == VTP_FIELD_Q ? s.Q : field == VTP_FIELD_VEL_GRAD ? s.L : s.stress;
            if (m.size() < n) return -2;
            for (uint32_t i = 0; i < n; ++i) for (int a = 0; a < 9; ++a) out[9 * i + a] = m[i].m[a];
            return 0;
        }
        // probes of the oracle's internal IMM-CT arrays (no device counterpart)
        case ORC_PROBE_BETA: return put1(out, s.beta_ct, n);
        case ORC_PROBE_KAPPA_CT: return put1(out, s.kappa_ct, n);
        case ORC_PROBE_BLOCKING: return put1(out, s.blocking, n);
        default: return -3;
    }
}

extern "C" void orc_set_stop(orc_sim *o, int stage) {
    if (o) o->s.stop_after = stage;
}

extern "C" int orc_write(orc_sim *o, int32_t field, const double *src) {
    if (!o || !src) return -1;
    Sim &s = o->s;
    if (field == VTP_FIELD_VEL_PHASE) {   // [n][P][3], after initialize
        if (s.uph.size() != (size_t)s.n * s.P) return -2;
        for (size_t i = 0; i < s.uph.size(); ++i) s.uph[i] = {src[3 * i], src[3 * i + 1], src[3 * i + 2]};
        return 0;
    }
    if (field == VTP_FIELD_Q) {           // [n][9] row-major, after initialize (IMM-CT)
        if (s.Q.size() != s.n) return -2;
        for (uint32_t i = 0; i < s.n; ++i)
            for (int a = 0; a < 9; ++a) s.Q[i].m[a] = src[9 * i + a];
        return 0;
    }
    std::vector<V3> *dst = field == VTP_FIELD_POS ? &s.pos : field == VTP_FIELD_VEL ? &s.vel : nullptr;
    if (!dst) return -3;
    for (uint32_t i = 0; i < s.n; ++i) (*dst)[i] = {src[3 * i], src[3 * i + 1], src[3 * i + 2]};
    if (field == VTP_FIELD_VEL && s.vadv.size() == s.n) s.vadv = s.vel;
    return 0;
}

// A neighbour search of the current positions alone (UGNSCudaImpl.cu:205-227),
// without a solver step: the per-tick search of a device run can be compared
// after its positions are written here.
extern "C" int orc_search(orc_sim *o) {
    if (!o) return -1;
    ugns_update(o->s);
    return 0;
}

extern "C" int orc_neighbors(orc_sim *o, uint32_t *lists, uint32_t *num, uint32_t *sorted, uint32_t *cell_start,
                             uint32_t *cell_end) {
    if (!o) return -1;
    Sim &s = o->s;
    if (lists) std::copy(s.nbrs.begin(), s.nbrs.end(), lists);
    if (num) std::copy(s.nnum.begin(), s.nnum.end(), num);
    if (sorted) std::copy(s.sorted.begin(), s.sorted.end(), sorted);
    size_t cells = (size_t)s.S * s.C;
    if (cell_start) std::copy(s.cstart.begin(), s.cstart.begin() + cells, cell_start);
    if (cell_end) std::copy(s.cend.begin(), s.cend.begin() + cells, cell_end);
    return 0;
}

extern "C" int orc_stats(orc_sim *o, uint32_t *div_it, uint32_t *inc_it, uint32_t *phase_it, double *div_res,
                         double *inc_res, uint32_t *crash) {
    if (!o) return -1;
    Sim &s = o->s;
    if (div_it) *div_it = s.div_it;
    if (inc_it) *inc_it = s.inc_it;
    if (phase_it) *phase_it = s.phase_it;
    if (div_res) *div_res = s.div_res;
    if (inc_res) *inc_res = s.inc_res;
    if (crash) *crash = s.crash;
    return 0;
}

extern "C" void orc_destroy(orc_sim *o) { delete o; }

// Kernel-function probes for the closed-form tests (KernelFunctions.cuh:13-69).
extern "C" void orc_kernel(double h, const double *r, double *w, double *grad, double *tension_c) {
    Kern k(h);
    V3 v{r[0], r[1], r[2]};
    if (w) *w = k.W(v);
    if (grad) {
        V3 g = k.grad(v);
        grad[0] = g.x;
        grad[1] = g.y;
        grad[2] = g.z;
    }
    if (tension_c) *tension_c = k.tension_C(length(v));
}

// The FP32 decision functions, exposed for the UGNS brute-force tests.
extern "C" int orc_cell_coord(float x, float lb, float cell) { return cell_coord(x, lb, cell); }

// One conformation update (row-major 3x3 arrays), for the closed-form pins.
extern "C" void orc_ct_update(const double *Q, const double *L, double dt, double relax, double beta, double kappa,
                              double vis_base, double vis_max, double a1, double th, double *Qout, double *mu,
                              double *stress) {
    M3 q, l, qo, st;
    for (int k = 0; k < 9; ++k) {
        q.m[k] = Q[k];
        l.m[k] = L[k];
    }
    ct_update(q, l, dt, relax, beta, kappa, vis_base, vis_max, a1, th, qo, *mu, st);
    for (int k = 0; k < 9; ++k) {
        Qout[k] = qo.m[k];
        stress[k] = st.m[k];
    }
}
extern "C" int orc_neighbor_test(const float *a, const float *b, float cell) {
    return neighbor_test(V3{a[0], a[1], a[2]}, V3{b[0], b[1], b[2]}, cell) ? 1 : 0;
}

#include "oracle_adjoint.inc"
