// vtp_device.h - device helpers shared by every solver: float3 arithmetic, the
// FP32 decision functions of the uniform-grid search (cell coordinate and the
// neighbour test, both bit-exact with the reference), the cubic-spline SPH
// kernels and wavefront reductions.  gfx950 only (wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// Build-time tuning of the neighbour loops (values measured on MI355X at
// PBF/DFSPH 1M particles):
//   VTP_NBR_BATCH      neighbours in flight per lane when a neighbour's data
//                      is at most 32 bytes (larger payloads use half)
//   VTP_PASS_MIN_WAVES minimum waves per SIMD the PBF neighbour passes are
//                      compiled for (bounds their registers)
#ifndef VTP_NBR_BATCH
#define VTP_NBR_BATCH 8
#endif
#ifndef VTP_PASS_MIN_WAVES
#define VTP_PASS_MIN_WAVES 1
#endif
//   VTP_NBR_LDS_DECODE 1: a lane's three layer bases sit in LDS, so decoding a
//                      16-bit entry is one LDS read instead of two selects
#ifndef VTP_NBR_LDS_DECODE
#define VTP_NBR_LDS_DECODE 1
#endif

namespace vtp {

// Launch-side description of the uniform grid, passed by value.
struct Grid {
    float lb_x, lb_y, lb_z;   // simSpaceLB
    float cell;               // gridCellSize == h (PBFSolver.cpp:89)
    int gx, gy, gz;           // ceil(simSpaceSize / cell) per axis
    uint32_t cells;           // gx*gy*gz, cells of one scene
    uint32_t scenes;          // independent scenes (batched mode)
    uint32_t dead_key;        // scenes*cells: sort key of particles off the grid
    float nbr_d2_max;         // largest float d2 with sqrtf(d2) <= cell (host-computed)
};

// Per-slot neighbour count word: low 31 bits = non-self neighbours stored,
// bit 31 = particle is off the grid ("dead", UGNSCudaImpl.cu:122-123).
constexpr uint32_t kDeadBit = 0x80000000u;
constexpr uint32_t kEmpty = 0xFFFFFFFFu;

__device__ __forceinline__ float3 xyz(float4 a) { return make_float3(a.x, a.y, a.z); }
__device__ __forceinline__ float3 add3(float3 a, float3 b) { return make_float3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 sub3(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 mul3(float3 a, float s) { return make_float3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3 fma3(float3 a, float s, float3 c) {
    return make_float3(fmaf(a.x, s, c.x), fmaf(a.y, s, c.y), fmaf(a.z, s, c.z));
}

// ---- decision functions (must match the oracle bit for bit) ---------------
// Cell coordinate along one axis: floor((x - lb) / cell) with IEEE-rounded
// subtraction and division (UGNSCudaImpl.cu:28-47).
__device__ __forceinline__ int cell_coord(float x, float lb, float cell) {
    return static_cast<int>(floorf(__fdiv_rn(__fsub_rn(x, lb), cell)));
}

// Sort key of a position: scene*cells + z*gy*gx + y*gx + x (UGNSCudaImpl.cu:58-63),
// or g.dead_key when the cell lies outside [0, grid) on any axis.  The
// reference accepts cp == gridSize (an inclusive bound, UGNSCudaImpl.cu:79-80)
// which aliases into the next row; here that cell is rejected as well.
__device__ __forceinline__ uint32_t cell_key(float3 p, uint32_t scene, const Grid &g) {
    int cx = cell_coord(p.x, g.lb_x, g.cell);
    int cy = cell_coord(p.y, g.lb_y, g.cell);
    int cz = cell_coord(p.z, g.lb_z, g.cell);
    bool in = cx >= 0 && cy >= 0 && cz >= 0 && cx < g.gx && cy < g.gy && cz < g.gz;
    if (!in) return g.dead_key;
    return scene * g.cells + (uint32_t)cz * (uint32_t)(g.gx * g.gy) + (uint32_t)cy * (uint32_t)g.gx + (uint32_t)cx;
}

// Squared distance with every operation rounded separately (no contraction):
// ((dx*dx + dy*dy) + dz*dz), the evaluation order of dot() in the reference.
// HIP's __fmul_rn / __fadd_rn are plain operators that the default
// -ffp-contract=fast-honor-pragmas fuses into FMAs, so the operations are
// written here under contract(off): the products stay rounded and pairs at
// exactly the cut-off are decided as the FP32 restatement decides them.
__device__ __forceinline__ float dist2_exact(float3 a, float3 b) {
#pragma clang fp contract(off)
    const float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    const float xx = dx * dx, yy = dy * dy, zz = dz * dz;
    return (xx + yy) + zz;
}

// Neighbour test of UGNSCudaImpl.cu:194: 0 < sqrtf(d2) <= cell.  sqrtf is
// correctly rounded and monotone, so {d2 : sqrtf(d2) <= cell} = [0, d2_max]
// where d2_max is found once on the host; sqrtf(d2) > 0 <=> d2 > 0.
__device__ __forceinline__ bool is_neighbor(float d2, float d2_max) {
    return d2 > 0.0f && d2 <= d2_max;
}

// ---- cubic spline kernel (KernelFunctions.cuh:13-56) ----------------------
// sigma = 8 / (pi h^3) with the reference's pi literal 3.14159265f.
struct Cubic {
    float h, inv_h, sigma;      // sigma = 8/(pi h^3)
    float grad_scale;           // sigma / h
    float w0;                   // W(0) = sigma
    float wt, wu;               // 2 sigma, -8 sigma: W = wt t^3 + wu u^3
    float gt, gu;               // -6 sigma/h, 24 sigma/h: |gradW| = gt t^2 + gu u^2
};

__host__ inline Cubic make_cubic(float h) {
    Cubic c;
    c.h = h;
    c.inv_h = 1.0f / h;
    const float pi = 3.14159265f;
    c.sigma = 8.0f / pi / (h * h * h);
    c.grad_scale = c.sigma / h;
    c.w0 = c.sigma;
    c.wt = 2.0f * c.sigma;
    c.wu = -8.0f * c.sigma;
    c.gt = -6.0f * c.grad_scale;
    c.gu = 24.0f * c.grad_scale;
    return c;
}

// Truncated-power form of the cubic spline.  With q = r/h, t = (1 - q)_+ and
// u = (1/2 - q)_+,
//   W(q)  = sigma (6q^3 - 6q^2 + 1) = 2 sigma (t^3 - 4 u^3)        (q <= 1/2)
//         = 2 sigma (1 - q)^3       = 2 sigma t^3                  (1/2 < q <= 1)
//   dW/dq = 6 sigma q (3q - 2)      = 2 sigma (-3 t^2 + 12 u^2)    (q <= 1/2)
//         = -6 sigma (1 - q)^2      = 2 sigma (-3 t^2)             (1/2 < q <= 1)
// and both vanish for q > 1, so one straight-line expression covers the
// reference's branches (KernelFunctions.cuh:13-56) with no compare or select.
// t and u are fma(-r, 1/h, c) clamped to [0, 1] (1 - q and 1/2 - q are at
// most 1 for q >= 0), which the compiler folds into the fma's clamp bit.
// r = r2 * rsqrt(r2): one transcendental (a correctly rounded sqrt costs a
// dozen instructions and changes W by < 1e-7 relative).
// Squared-distance bound of the reference's q < 1e-6 gradient cut-off.
__device__ __forceinline__ float cubic_r2_lo(const Cubic &k) { return (1e-6f * k.h) * (1e-6f * k.h); }

// rsqrt of max(r2, FLT_MIN): finite for a coincident pair, whose r is then 0
// (W = sigma), so no neighbour loop branches on r2 == 0.  Identical to
// rsqrt(r2) for every normal r2.
__device__ __forceinline__ float cubic_rinv(float r2) { return __frsqrt_rn(fmaxf(r2, 1.17549435e-38f)); }

__device__ __forceinline__ float unit_clamp(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }

__device__ __forceinline__ float cubic_w(float r2, const Cubic &k) {
    const float r = r2 * cubic_rinv(r2);
    const float t = unit_clamp(fmaf(-r, k.inv_h, 1.0f)), u = unit_clamp(fmaf(-r, k.inv_h, 0.5f));
    return fmaf((t * t) * t, k.wt, ((u * u) * u) * k.wu);
}

// Scalar factor f such that gradW(r_vec) = f * r_vec (zero for q < 1e-6).
__device__ __forceinline__ float cubic_grad_factor(float r2, const Cubic &k) {
    const float rinv = cubic_rinv(r2);
    const float r = r2 * rinv;
    const float t = unit_clamp(fmaf(-r, k.inv_h, 1.0f)), u = unit_clamp(fmaf(-r, k.inv_h, 0.5f));
    const float f = fmaf(t * t, k.gt, (u * u) * k.gu) * rinv;
    return r2 >= cubic_r2_lo(k) ? f : 0.0f;
}

// W and the gradient factor together (one rsqrt), straight-line code.
__device__ __forceinline__ void cubic_w_grad(float r2, const Cubic &k, float &w, float &f) {
    const float rinv = cubic_rinv(r2);
    const float r = r2 * rinv;
    const float t = unit_clamp(fmaf(-r, k.inv_h, 1.0f)), u = unit_clamp(fmaf(-r, k.inv_h, 0.5f));
    const float t2 = t * t, u2 = u * u;
    w = fmaf(t2 * t, k.wt, (u2 * u) * k.wu);   // r2 == 0: r = 0, W = sigma
    const float g = fmaf(t2, k.gt, u2 * k.gu) * rinv;
    f = r2 >= cubic_r2_lo(k) ? g : 0.0f;
}

// Akinci cohesion spline C(r) (KernelFunctions.cuh:58-69), c = 32/(pi h^9).
__device__ __forceinline__ float surface_tension_c(float r, float h, float c_sigma, float h6_64) {
    float hr = h - r;
    float v = hr * hr * hr * r * r * r;
    if (r * 2.0f > h && r <= h) return c_sigma * v;
    if (r > 0.0f && r * 2.0f <= h) return c_sigma * (2.0f * v - h6_64);
    return 0.0f;
}

// ---- XCD-aware slot mapping ---------------------------------------------------
// The dispatcher deals workgroups round-robin over the 8 XCDs of the die
// (workgroup b runs on XCD b % 8), and each XCD has its own L2.  Slots are in
// cell order, so giving every XCD one contiguous eighth of the workgroups
// keeps a particle's neighbours (and the neighbours of its neighbours) in the
// same L2 instead of spreading every neighbourhood over all eight.
constexpr uint32_t kXcds = 8;
constexpr uint32_t kMaxWavesPerBlock = 16;   // 1024-thread workgroups at most
__device__ __forceinline__ uint32_t xcd_block() {
    const uint32_t b = blockIdx.x, nb = gridDim.x;
    const uint32_t x = b % kXcds, k = b / kXcds;   // XCD, rank inside that XCD
    const uint32_t q = nb / kXcds, r = nb % kXcds; // the first r XCDs get q + 1 blocks
    return x * q + min(x, r) + k;
}
__device__ __forceinline__ uint32_t slot_index() { return xcd_block() * blockDim.x + threadIdx.x; }

// ---- neighbour iteration ---------------------------------------------------
// Gathers through a 32-bit byte offset from a wave-uniform base: the loads
// take the SGPR-base + VGPR-offset form, with no 64-bit address arithmetic per
// neighbour.  Needs n * sizeof(T) < 4 GiB (vtp_initialize caps n at 2^28).
template <typename T>
__device__ __forceinline__ T ldg(const T *__restrict__ base, uint32_t j) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + (uint64_t)(j * (uint32_t)sizeof(T)));
}
// Only the w component of a float4 element (a 4-byte load).
__device__ __forceinline__ float ldg_w(const float4 *__restrict__ base, uint32_t j) {
    return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(base) + (uint64_t)(j * 16u + 12u));
}

// ---- neighbour lists -------------------------------------------------------
// One allocation per search (Ugns::nbr), addressed through its first word:
//   uint32_t base[3][stride]       per cell layer dz = -1, 0, +1 of the
//                                  particle's 27 cells: the lowest slot of the
//                                  layer's 9 cells (the start of its first
//                                  non-empty cell row in key order)
//   uint32_t wide[flag words]      one word per 64-slot group: 1 = that group's
//                                  lists hold full 32-bit entries
//   16-byte chunks chunk[q][stride] chunk q of slot i at chunk index
//                                  q * stride + i: a wavefront reads 1 KB of
//                                  consecutive chunks.  In a narrow group
//                                  chunk q holds entries 8q .. 8q + 7 as
//                                  16-bit e, entry = base[e >> 14][i] +
//                                  (e & 0x3FFF); in a wide group it holds
//                                  entries 4q .. 4q + 3 as 32-bit slots.
// Keys are z-major, so the 9 cells of one layer span three cell rows and a
// neighbour lies at most two rows (~2 gx cells) past its layer's base: 14
// bits hold that unless a cell row carries more than ~5k particles, and then
// only that particle's 64-slot group switches to 32-bit entries.  A pass
// fetches eight entries with one 16-byte load, so its index loads are one
// vector-memory instruction per eight neighbours, not one per neighbour.
// Entries past a list's length in its last chunk are stale; readers never
// use them.  stride is a multiple of 64; the flag words are padded so that
// the chunks start on a 256-byte boundary.
constexpr uint32_t kNbrOffBits = 14;
constexpr uint32_t kNbrOffMask = (1u << kNbrOffBits) - 1u;
__host__ __device__ inline uint32_t nbr_flag_words(uint32_t stride) { return ((stride / 64u + 63u) / 64u) * 64u; }
// chunks per slot: enough for cap 32-bit entries, rounded to whole 8-entry
// batches (two chunks each in a wide group)
__host__ __device__ inline uint32_t nbr_chunks(uint32_t cap) { return 2u * ((cap + 7u) / 8u); }
__host__ __device__ inline size_t nbr_words(uint32_t cap, uint32_t stride) {
    return 3 * (size_t)stride + nbr_flag_words(stride) + (size_t)nbr_chunks(cap) * 4u * stride;
}
__host__ __device__ inline uint4 *nbr_chunks_of(uint32_t *nbr, uint32_t stride) {
    return reinterpret_cast<uint4 *>(nbr + 3 * (size_t)stride + nbr_flag_words(stride));
}
__host__ __device__ inline const uint4 *nbr_chunks_of(const uint32_t *nbr, uint32_t stride) {
    return reinterpret_cast<const uint4 *>(nbr + 3 * (size_t)stride + nbr_flag_words(stride));
}
__host__ __device__ inline uint32_t nbr_flag_index(uint32_t stride, uint32_t i) { return 3 * stride + (i >> 6); }
// Element positions of entry t of slot i: a 16-bit index into the chunks for
// a narrow entry, a 32-bit index for a wide one.
__host__ __device__ inline size_t nbr_narrow_pos(uint32_t stride, uint32_t i, uint32_t t) {
    return ((size_t)(t >> 3) * stride + i) * 8u + (t & 7u);
}
__host__ __device__ inline size_t nbr_wide_pos(uint32_t stride, uint32_t i, uint32_t t) {
    return ((size_t)(t >> 2) * stride + i) * 4u + (t & 3u);
}
__device__ __forceinline__ uint32_t nbr_decode(uint32_t e, uint32_t b0, uint32_t b1, uint32_t b2) {
    const uint32_t z = e >> kNbrOffBits;
    return (z == 0 ? b0 : z == 1 ? b1 : b2) + (e & kNbrOffMask);
}
// Entry t of slot i's list, decoded per lane (for kernels off the hot path).
__device__ __forceinline__ uint32_t nbr_entry(const uint32_t *__restrict__ nbr, uint32_t stride, uint32_t i,
                                              uint32_t t) {
    const uint4 *ch = nbr_chunks_of(nbr, stride);
    if (nbr[nbr_flag_index(stride, i)]) return reinterpret_cast<const uint32_t *>(ch)[nbr_wide_pos(stride, i, t)];
    return nbr_decode(reinterpret_cast<const uint16_t *>(ch)[nbr_narrow_pos(stride, i, t)], nbr[i], nbr[stride + i],
                      nbr[2 * (size_t)stride + i]);
}

// Chunk readers of one slot's list.  `row` is the byte offset of the slot's
// chunk 0 (i * 16); chunk q is `step` bytes further on per q.  batch<B>(t, j)
// fills j[0 .. B) with entries t .. t + B - 1 (t a multiple of B, B = 4 or 8)
// from one or two 16-byte loads; entry(t) reads one entry.
struct NarrowIdx {
    const char *ch;
    size_t step;
    uint32_t row, b0, b1, b2;
    __device__ __forceinline__ uint32_t entry(uint32_t t) const {
        const char *p = ch + (size_t)(t >> 3) * step;
        return nbr_decode(*reinterpret_cast<const uint16_t *>(p + (row + (t & 7u) * 2u)), b0, b1, b2);
    }
    template <uint32_t B>
    __device__ __forceinline__ void batch(uint32_t t, uint32_t *j) const {
        const char *p = ch + (size_t)(t >> 3) * step;
        if (B == 8) {
            const uint4 w = *reinterpret_cast<const uint4 *>(p + row);
            const uint32_t v[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
            for (uint32_t k = 0; k < 4; ++k) {
                j[2 * k] = nbr_decode(v[k] & 0xFFFFu, b0, b1, b2);
                j[2 * k + 1] = nbr_decode(v[k] >> 16, b0, b1, b2);
            }
        } else {
            const uint2 w = *reinterpret_cast<const uint2 *>(p + (row + (t & 7u) * 2u));
            j[0] = nbr_decode(w.x & 0xFFFFu, b0, b1, b2);
            j[1] = nbr_decode(w.x >> 16, b0, b1, b2);
            j[2] = nbr_decode(w.y & 0xFFFFu, b0, b1, b2);
            j[3] = nbr_decode(w.y >> 16, b0, b1, b2);
        }
    }
};
// Workgroup size up to which the neighbour passes keep their layer bases in
// LDS (3 KB per workgroup); every tuned pass shape is 64, 128 or 256.
constexpr uint32_t kNbrLdsLanes = 256;
// The same reader with the lane's layer bases in LDS, word z * kNbrLdsLanes
// + threadIdx.x for layer z (consecutive lanes, consecutive banks; the fixed
// row length makes the address one shift-add).  Each lane writes and reads
// only its own words, so no barrier is needed.
struct NarrowIdxLds {
    const char *ch;
    size_t step;
    uint32_t row;
    const uint32_t *lds;   // this lane's word of layer 0
    __device__ __forceinline__ uint32_t dec(uint32_t e) const {
        return lds[(e >> kNbrOffBits) * kNbrLdsLanes] + (e & kNbrOffMask);
    }
    __device__ __forceinline__ uint32_t entry(uint32_t t) const {
        const char *p = ch + (size_t)(t >> 3) * step;
        return dec(*reinterpret_cast<const uint16_t *>(p + (row + (t & 7u) * 2u)));
    }
    template <uint32_t B>
    __device__ __forceinline__ void batch(uint32_t t, uint32_t *j) const {
        const char *p = ch + (size_t)(t >> 3) * step;
        if (B == 8) {
            const uint4 w = *reinterpret_cast<const uint4 *>(p + row);
            const uint32_t v[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
            for (uint32_t k = 0; k < 4; ++k) {
                j[2 * k] = dec(v[k] & 0xFFFFu);
                j[2 * k + 1] = dec(v[k] >> 16);
            }
        } else {
            const uint2 w = *reinterpret_cast<const uint2 *>(p + (row + (t & 7u) * 2u));
            j[0] = dec(w.x & 0xFFFFu);
            j[1] = dec(w.x >> 16);
            j[2] = dec(w.y & 0xFFFFu);
            j[3] = dec(w.y >> 16);
        }
    }
};
struct WideIdx {
    const char *ch;
    size_t step;
    uint32_t row;
    __device__ __forceinline__ uint32_t entry(uint32_t t) const {
        const char *p = ch + (size_t)(t >> 2) * step;
        return *reinterpret_cast<const uint32_t *>(p + (row + (t & 3u) * 4u));
    }
    template <uint32_t B>
    __device__ __forceinline__ void batch(uint32_t t, uint32_t *j) const {
#pragma unroll
        for (uint32_t h = 0; h < B / 4; ++h) {
            const uint4 w = *reinterpret_cast<const uint4 *>(ch + (size_t)((t >> 2) + h) * step + row);
            j[4 * h] = w.x;
            j[4 * h + 1] = w.y;
            j[4 * h + 2] = w.z;
            j[4 * h + 3] = w.w;
        }
    }
};

// Walks slot i's list in batches: the gathers of a batch are issued
// together, and the entries of the next batch are requested before the
// current batch is used, so the index -> gather dependency of one batch
// overlaps the arithmetic of the previous one.  A batch is 8 neighbours when
// a neighbour's data fits in 32 bytes and 4 when it is larger (register
// pressure).  The last batch may be partial: its stale entries are replaced
// by `self` (a valid slot) for the gathers and not used.
template <typename Idx, typename Load, typename Use>
__device__ __forceinline__ void walk_list(const Idx &idx, uint32_t self, uint32_t nn, Load &&load, Use &&use) {
    using D = decltype(load(0u));
    constexpr uint32_t B = sizeof(D) > 32 ? VTP_NBR_BATCH / 2 : VTP_NBR_BATCH;
    static_assert(B == 4 || B == 8, "a batch is half or all of an 8-entry chunk");
    if (nn == 0) return;
    uint32_t j[B];
    idx.template batch<B>(0, j);
    uint32_t t = 0;
    for (; t + B <= nn; t += B) {
        D d[B];
#pragma unroll
        for (uint32_t b = 0; b < B; ++b) d[b] = load(j[b]);
        if (t + B < nn) idx.template batch<B>(t + B, j);
#pragma unroll
        for (uint32_t b = 0; b < B; ++b) use(d[b]);
    }
    if (t < nn) {
        const uint32_t rem = nn - t;
        D d[B];
#pragma unroll
        for (uint32_t b = 0; b < B; ++b) d[b] = load(b < rem ? j[b] : self);
#pragma unroll
        for (uint32_t b = 0; b < B; ++b)
            if (b < rem) use(d[b]);
    }
}

// The list of slot i, nn entries.  The format flag belongs to i's 64-slot
// group, which is the same in every lane of a wavefront here (waves cover
// 64 consecutive, 64-aligned slots, or 64/L particles with L lanes each), so
// it is read once per wave and the branch is uniform.
template <typename Load, typename Use>
__device__ __forceinline__ void for_each_neighbor(const uint32_t *__restrict__ nbr, uint32_t stride, uint32_t i,
                                                  uint32_t nn, Load &&load, Use &&use) {
    const char *ch = reinterpret_cast<const char *>(nbr_chunks_of(nbr, stride));
    const size_t step = (size_t)stride * sizeof(uint4);
    const uint32_t row = i * (uint32_t)sizeof(uint4);
    if (__builtin_amdgcn_readfirstlane(nbr[nbr_flag_index(stride, i)]))
        walk_list(WideIdx{ch, step, row}, i, nn, load, use);
    else
#if VTP_NBR_LDS_DECODE
    if (blockDim.x <= kNbrLdsLanes) {
        __shared__ uint32_t s_bases[3 * kNbrLdsLanes];
        uint32_t *mine = s_bases + threadIdx.x;
        mine[0] = nbr[i];
        mine[kNbrLdsLanes] = nbr[stride + i];
        mine[2 * kNbrLdsLanes] = nbr[2 * (size_t)stride + i];
        walk_list(NarrowIdxLds{ch, step, row, mine}, i, nn, load, use);
    } else
#endif
        walk_list(NarrowIdx{ch, step, row, nbr[i], nbr[stride + i], nbr[2 * (size_t)stride + i]}, i, nn, load, use);
}

// Cooperative walk: L lanes share one particle, lane `sub` taking entries
// sub, sub + L, sub + 2L, ...  In one gather instruction the L lanes of a
// particle then read consecutive entries of its list (mostly neighbouring
// slots of one cell row), so a wave touches far fewer cache lines than with
// one particle per lane; the partial sums are combined with lane_sum<L>.
template <uint32_t L, typename Idx, typename Load, typename Use>
__device__ __forceinline__ void walk_list_coop(const Idx &idx, uint32_t nn, uint32_t sub, Load &&load, Use &&use) {
    using D = decltype(load(0u));
    constexpr uint32_t B = sizeof(D) > 32 ? 2 : 4;   // neighbours in flight per lane
    uint32_t t = sub;
    for (; t + (B - 1) * L < nn; t += B * L) {
        uint32_t j[B];
#pragma unroll
        for (uint32_t b = 0; b < B; ++b) j[b] = idx.entry(t + b * L);
        D d[B];
#pragma unroll
        for (uint32_t b = 0; b < B; ++b) d[b] = load(j[b]);
#pragma unroll
        for (uint32_t b = 0; b < B; ++b) use(d[b]);
    }
    for (; t < nn; t += L) use(load(idx.entry(t)));
}

template <uint32_t L, typename Load, typename Use>
__device__ __forceinline__ void for_each_neighbor_coop(const uint32_t *__restrict__ nbr, uint32_t stride,
                                                       uint32_t i, uint32_t nn, uint32_t sub, Load &&load,
                                                       Use &&use) {
    if (L == 1) {
        for_each_neighbor(nbr, stride, i, nn, load, use);
        return;
    }
    const char *ch = reinterpret_cast<const char *>(nbr_chunks_of(nbr, stride));
    const size_t step = (size_t)stride * sizeof(uint4);
    const uint32_t row = i * (uint32_t)sizeof(uint4);
    if (__builtin_amdgcn_readfirstlane(nbr[nbr_flag_index(stride, i)]))
        walk_list_coop<L>(WideIdx{ch, step, row}, nn, sub, load, use);
    else
        walk_list_coop<L>(NarrowIdx{ch, step, row, nbr[i], nbr[stride + i], nbr[2 * (size_t)stride + i]}, nn, sub,
                          load, use);
}

// Sum over the L lanes (L = 1, 2 or 4, aligned groups) sharing one particle.
template <uint32_t L>
__device__ __forceinline__ float lane_sum(float v) {
    if (L >= 2) v += __shfl_xor(v, 1, 64);
    if (L >= 4) v += __shfl_xor(v, 2, 64);
    return v;
}

// ---- wavefront reductions (wave64) ----------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, off, 64));
    return v;
}
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

}  // namespace vtp
