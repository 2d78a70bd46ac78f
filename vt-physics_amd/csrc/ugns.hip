// ugns.hip - uniform-grid neighbour search for MI355X.
//
// Replaces UGNSCudaImpl.cu:96-227 (resetDevPtr, calcParticleHashValue,
// sortByHashValue = 2x thrust::sort_by_key, findCellRange, findNeighbors).
//
// Differences in structure (results are identical):
//  * particle state is kept in cell-sorted slot order and physically
//    reordered after every sort, so neighbour gathers hit contiguous memory;
//  * particles are grouped by cell with a counting sort written here (one
//    atomic per run of equal keys in a wavefront, a scan over the cells that
//    yields cellStart/cellEnd directly, a scatter); a stable LSD radix sort
//    (8-bit digits, a per-tile digit histogram, a per-digit spine scan and a
//    scatter whose stable in-tile ranks come from wavefront ballots: one
//    64-bit ballot per key bit gives every lane the mask of lanes holding the
//    same digit) is kept as an option and gives the same order;
//  * because the input is the previous step's slot order rather than attach
//    order, equal cell ids are put in attach-id order by a per-cell rank pass,
//    which reproduces thrust's stable (cellId, attachId) order exactly;
//  * neighbour lists hold 16-bit slot offsets from a per-particle base,
//    eight to a 16-byte chunk, chunk q of slot i at chunk q * stride + i
//    (lane i of a wavefront reads chunk i: 1 KB, fully coalesced; layout in
//    vtp_device.h), without the self entry; the reference layout is rebuilt
//    on request.
#include <algorithm>

#include "vtp_internal.h"

namespace vtp {

constexpr int kRsThreads = 256;
constexpr int kRsRounds = 4;                    // keys per lane per tile
constexpr int kRsWaveChunk = 64 * kRsRounds;    // 256 keys per wavefront
constexpr int kRsTile = kRsThreads * kRsRounds; // 1024 keys per workgroup (1.3k tiles at 1.3M keys)
// Neighbour build: 128 threads, each with a private LDS column of up to
// kNbrStage 16-bit entries.
constexpr uint32_t kNbrBuildThreads = 128;
#ifndef VTP_NBR_STAGE
#define VTP_NBR_STAGE 32
#endif
#ifndef VTP_BUILD_BATCH
#define VTP_BUILD_BATCH 8
#endif
#ifndef VTP_NBR_STAGE_LAMBDA
#define VTP_NBR_STAGE_LAMBDA 48
#endif
// At least 6 waves per SIMD: caps the build at 80 VGPRs.  The plain build
// needs 59 (8 waves); the one that also sums lambda would take 81 (5 waves)
// and instead spills one register.
#ifndef VTP_BUILD_MIN_WAVES
#define VTP_BUILD_MIN_WAVES 6
#endif
#ifndef VTP_BUILD_BATCH_LAMBDA
#define VTP_BUILD_BATCH_LAMBDA 8
#endif
#ifndef VTP_REF_BATCH
#define VTP_REF_BATCH 8
#endif
#ifndef VTP_REF_ORDER_SHARE
#define VTP_REF_ORDER_SHARE 8
#endif
constexpr uint32_t kNbrStage = VTP_NBR_STAGE;   // list entries per thread staged in LDS (8 KB per workgroup)
static_assert(VTP_NBR_STAGE % 8 == 0 && VTP_NBR_STAGE_LAMBDA % 8 == 0, "the stage is flushed in whole 8-entry chunks");
// The PBF build that also computes lambda needs more registers (6 waves per
// SIMD), which leaves room for a larger stage at the same occupancy.
constexpr uint32_t kNbrStageLambda = VTP_NBR_STAGE_LAMBDA;   // 12 KB per workgroup
// A build walks every particle in the reference's cell order when more than
// 1/kRefOrderShare of the particles reached the cap in the previous build.
constexpr uint32_t kRefOrderShare = VTP_REF_ORDER_SHARE;

__global__ void k_fill_u32(uint32_t *a, uint32_t n, uint32_t v) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// Digit histogram of one tile.  Keys of consecutive slots usually share a
// digit, so lanes first merge equal digits with ballots and only the leader
// of each group touches LDS.
__global__ __launch_bounds__(kRsThreads) void k_radix_hist(const uint32_t *__restrict__ keys, uint32_t n,
                                                           int shift, uint32_t *__restrict__ spine,
                                                           uint32_t tiles) {
    __shared__ uint32_t h[256];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    h[tid] = 0;
    __syncthreads();
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint32_t base = blockIdx.x * kRsTile + w * kRsWaveChunk;
    for (int r = 0; r < kRsRounds; ++r) {
        uint32_t i = base + r * 64 + lane;
        bool valid = i < n;
        uint32_t d = valid ? (keys[i] >> shift) & 255u : 0u;
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            uint64_t bb = __ballot((d >> b) & 1u);
            m &= ((d >> b) & 1u) ? bb : ~bb;
        }
        if (valid && (m & lt) == 0) atomicAdd(&h[d], (uint32_t)__popcll(m));
    }
    __syncthreads();
    spine[(uint32_t)tid * tiles + blockIdx.x] = h[tid];
}

// One workgroup per digit: exclusive scan of that digit's per-tile counts,
// total into dtot[d].
__global__ __launch_bounds__(kRsThreads) void k_radix_spine(uint32_t *__restrict__ spine, uint32_t tiles,
                                                            uint32_t *__restrict__ dtot) {
    __shared__ uint32_t s[kRsThreads];
    const int tid = threadIdx.x;
    uint32_t *row = spine + (uint32_t)blockIdx.x * tiles;
    const uint32_t per = (tiles + kRsThreads - 1) / kRsThreads;
    const uint32_t b = tid * per, e = min(b + per, tiles);
    uint32_t sum = 0;
    for (uint32_t t = b; t < e; ++t) sum += row[t];
    s[tid] = sum;
    __syncthreads();
    for (int off = 1; off < kRsThreads; off <<= 1) {
        uint32_t v = tid >= off ? s[tid - off] : 0u;
        __syncthreads();
        s[tid] += v;
        __syncthreads();
    }
    uint32_t run = s[tid] - sum;  // exclusive prefix of this thread's chunk
    for (uint32_t t = b; t < e; ++t) {
        uint32_t c = row[t];
        row[t] = run;
        run += c;
    }
    if (tid == kRsThreads - 1) dtot[blockIdx.x] = s[tid];
}

// Stable scatter of one tile.  Wavefront w owns keys [w*256, (w+1)*256) of
// the tile, processed in 4 rounds of 64 in index order, so the order
// (wave, round, lane) equals the input order and ranks are stable.
__global__ __launch_bounds__(kRsThreads) void k_radix_scatter(const uint32_t *__restrict__ kin,
                                                              const uint32_t *__restrict__ vin,
                                                              uint32_t *__restrict__ kout,
                                                              uint32_t *__restrict__ vout, uint32_t n,
                                                              int shift, const uint32_t *__restrict__ spine,
                                                              const uint32_t *__restrict__ dtot,
                                                              uint32_t tiles) {
    __shared__ uint32_t wc[4][256];   // per-wave running digit counts
    __shared__ uint32_t toff[256];    // global offset of each digit for this tile
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;

    // exclusive scan of the 256 digit totals
    toff[tid] = dtot[tid];
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        uint32_t v = tid >= off ? toff[tid - off] : 0u;
        __syncthreads();
        toff[tid] += v;
        __syncthreads();
    }
    uint32_t incl = toff[tid];
    __syncthreads();
    toff[tid] = incl - dtot[tid] + spine[(uint32_t)tid * tiles + blockIdx.x];
#pragma unroll
    for (int q = 0; q < 4; ++q) wc[q][tid] = 0;
    __syncthreads();

    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const uint32_t base = blockIdx.x * kRsTile + w * kRsWaveChunk;
    uint32_t kr[kRsRounds], vr[kRsRounds], rr[kRsRounds];
#pragma unroll
    for (int r = 0; r < kRsRounds; ++r) {
        uint32_t i = base + r * 64 + lane;
        bool valid = i < n;
        uint32_t k = valid ? kin[i] : 0u;
        uint32_t d = (k >> shift) & 255u;
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            uint64_t bb = __ballot((d >> b) & 1u);
            m &= ((d >> b) & 1u) ? bb : ~bb;
        }
        uint32_t rank = (uint32_t)__popcll(m & lt);
        uint32_t prev = wc[w][d];
        // the leader's store depends on `prev`, so every lane's read of this
        // round completes before the running count moves on
        if (valid && rank == 0) wc[w][d] = prev + (uint32_t)__popcll(m);
        kr[r] = k;
        vr[r] = valid ? vin[i] : 0u;
        rr[r] = prev + rank;
    }
    __syncthreads();
    {   // per digit: exclusive prefix over the four wavefronts
        uint32_t s = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            uint32_t t = wc[q][tid];
            wc[q][tid] = s;
            s += t;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRsRounds; ++r) {
        uint32_t i = base + r * 64 + lane;
        if (i < n) {
            uint32_t d = (kr[r] >> shift) & 255u;
            uint32_t pos = toff[d] + wc[w][d] + rr[r];
            kout[pos] = kr[r];
            vout[pos] = vr[r];
        }
    }
}

// ---------------------------------------------------------------- counting sort
// The in-cell order is fixed afterwards by k_cell_fixup (attach id), so the
// grouping by cell needs no stable order for live particles: every particle
// takes a rank in its cell from one atomic per run of equal keys, a scan over
// the cells gives cellStart / cellEnd, and a scatter places the particle at
// start + rank.  Particles off the grid keep the previous slot order (as the
// stable radix sort does): their ranks come from per-tile counts.
constexpr uint32_t kCsThreads = 256;         // one particle per thread, tile = workgroup
constexpr uint32_t kScanPer = 4;             // cells per thread in the cell scan
constexpr uint32_t kScanBlock = kCsThreads * kScanPer;

__device__ __forceinline__ uint64_t lanes_below(int lane) { return lane == 0 ? 0ull : (~0ull >> (64 - lane)); }

// Exclusive prefix of the per-wave values v[w] (w < 4) for wave w, via LDS.
__device__ __forceinline__ uint32_t wave_offset(uint32_t *s4, int w) {
    uint32_t o = 0;
    for (int q = 0; q < w; ++q) o += s4[q];
    return o;
}

__global__ __launch_bounds__(kCsThreads) void k_cell_count(const uint32_t *__restrict__ key, uint32_t n,
                                                           uint32_t dead_key, uint32_t *__restrict__ count,
                                                           uint32_t *__restrict__ rank,
                                                           uint32_t *__restrict__ tile_dead) {
    __shared__ uint32_t s_dead[kCsThreads / 64];
    const uint32_t i = blockIdx.x * kCsThreads + threadIdx.x;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const bool valid = i < n;
    const uint32_t k = valid ? key[i] : kEmpty;
    const bool live = valid && k < dead_key;
    // runs of equal live keys in consecutive lanes (slots are cell-sorted
    // from the previous search, so a cell is mostly one run)
    const uint32_t kprev = __shfl_up(k, 1, 64);
    const bool head = live && (lane == 0 || kprev != k);
    const uint64_t heads = __ballot(head);
    const uint64_t stops = heads | __ballot(!live);   // a run ends before the next head or non-live lane
    const int h = live ? 63 - __clzll(heads & (~0ull >> (63 - lane))) : lane;
    uint32_t base = 0;
    if (head) {
        const uint64_t above = stops & ~(~0ull >> (63 - lane));   // lanes > this one
        const int next = above ? __ffsll((long long)above) - 1 : 64;
        base = atomicAdd(&count[k], (uint32_t)(next - lane));
    }
    base = __shfl(base, h, 64);
    if (live) rank[i] = base + (uint32_t)(lane - h);
    const uint64_t dm = __ballot(valid && !live);
    if (lane == 0) s_dead[w] = (uint32_t)__popcll(dm);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int q = 0; q < kCsThreads / 64; ++q) t += s_dead[q];
        tile_dead[blockIdx.x] = t;
    }
}

// Sum of kScanBlock cell counts per workgroup.
__global__ __launch_bounds__(kCsThreads) void k_cell_scan_reduce(const uint32_t *__restrict__ count, uint32_t ncell,
                                                                 uint32_t *__restrict__ bsum) {
    __shared__ uint32_t s4[kCsThreads / 64];
    const uint32_t b0 = blockIdx.x * kScanBlock + threadIdx.x * kScanPer;
    uint32_t t = 0;
#pragma unroll
    for (uint32_t q = 0; q < kScanPer; ++q) t += b0 + q < ncell ? count[b0 + q] : 0u;
    t = wave_sum_u32(t);
    if ((threadIdx.x & 63) == 0) s4[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) bsum[blockIdx.x] = s4[0] + s4[1] + s4[2] + s4[3];
}

// Exclusive scan of a[0..m) in place by one workgroup of kSpineThreads, in
// rounds of kSpineThreads elements (wave scans, then the wave totals);
// returns the total.
constexpr uint32_t kSpineThreads = 1024;
__device__ uint32_t workgroup_exclusive_scan(uint32_t *a, uint32_t m, uint32_t *s_w) {
    const uint32_t tid = threadIdx.x;
    const int lane = tid & 63, w = tid >> 6;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < m; base += kSpineThreads) {
        const uint32_t idx = base + tid;
        const uint32_t v = idx < m ? a[idx] : 0u;
        uint32_t incl = v;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t u = __shfl_up(incl, off, 64);
            if (lane >= off) incl += u;
        }
        if (lane == 63) s_w[w] = incl;
        __syncthreads();
        uint32_t woff = 0, rtot = 0;
#pragma unroll
        for (int q = 0; q < (int)(kSpineThreads / 64); ++q) {
            const uint32_t t = s_w[q];
            woff += q < w ? t : 0u;
            rtot += t;
        }
        if (idx < m) a[idx] = carry + woff + incl - v;
        carry += rtot;
        __syncthreads();   // s_w is rewritten by the next round
    }
    return carry;
}

// One workgroup: block offsets of the cell scan, tile offsets of the
// off-grid particles (only when there are any), and the dead key's range
// (after every live particle).
__global__ __launch_bounds__(kSpineThreads) void k_cell_scan_spine(uint32_t *__restrict__ bsum, uint32_t nblocks,
                                                                   uint32_t *__restrict__ tile_dead, uint32_t tiles,
                                                                   uint32_t n, uint32_t dead_key,
                                                                   uint32_t *__restrict__ start,
                                                                   uint32_t *__restrict__ end,
                                                                   uint32_t *__restrict__ meta) {
    __shared__ uint32_t s_w[kSpineThreads / 64];
    const uint32_t live = workgroup_exclusive_scan(bsum, nblocks, s_w);
    const uint32_t dead = n - live;
    if (dead) workgroup_exclusive_scan(tile_dead, tiles, s_w);
    if (threadIdx.x == 0) {
        meta[0] = live;
        start[dead_key] = dead ? live : kEmpty;
        end[dead_key] = dead ? n - 1 : kEmpty;
    }
}

// cellStart / cellEnd (inclusive, kEmpty for an empty cell) from the counts;
// the counts are cleared for the next search.
__global__ __launch_bounds__(kCsThreads) void k_cell_scan_ranges(uint32_t *__restrict__ count, uint32_t ncell,
                                                                 const uint32_t *__restrict__ bsum,
                                                                 uint32_t *__restrict__ start,
                                                                 uint32_t *__restrict__ end) {
    __shared__ uint32_t s4[kCsThreads / 64];
    const uint32_t b0 = blockIdx.x * kScanBlock + threadIdx.x * kScanPer;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t c[kScanPer], t = 0;
#pragma unroll
    for (uint32_t q = 0; q < kScanPer; ++q) {
        c[q] = b0 + q < ncell ? count[b0 + q] : 0u;
        t += c[q];
    }
    // exclusive prefix of t over the workgroup: in-wave scan, then wave offsets
    uint32_t incl = t;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t v = __shfl_up(incl, off, 64);
        if (lane >= off) incl += v;
    }
    if (lane == 63) s4[w] = incl;
    __syncthreads();
    uint32_t run = bsum[blockIdx.x] + wave_offset(s4, w) + incl - t;
#pragma unroll
    for (uint32_t q = 0; q < kScanPer; ++q) {
        if (b0 + q < ncell) {
            start[b0 + q] = c[q] ? run : kEmpty;
            end[b0 + q] = c[q] ? run + c[q] - 1 : kEmpty;
            if (c[q]) count[b0 + q] = 0u;
        }
        run += c[q];
    }
}

// Places every particle: live ones at start[key] + rank, off-grid ones after
// the live particles in previous-slot order.  Also writes the sorted key, the
// old slot and its attach id.
__global__ __launch_bounds__(kCsThreads) void k_cell_scatter(const uint32_t *__restrict__ key, uint32_t n,
                                                             uint32_t dead_key, const uint32_t *__restrict__ rank,
                                                             const uint32_t *__restrict__ start,
                                                             const uint32_t *__restrict__ tile_dead,
                                                             const uint32_t *__restrict__ meta,
                                                             const uint32_t *__restrict__ orig_old,
                                                             uint32_t *__restrict__ kout, uint32_t *__restrict__ vout,
                                                             uint32_t *__restrict__ so) {
    __shared__ uint32_t s_dead[kCsThreads / 64];
    const uint32_t i = blockIdx.x * kCsThreads + threadIdx.x;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const bool valid = i < n;
    const uint32_t k = valid ? key[i] : kEmpty;
    const bool dead = valid && k >= dead_key;
    const uint64_t dm = __ballot(dead);
    if (lane == 0) s_dead[w] = (uint32_t)__popcll(dm);
    __syncthreads();
    if (!valid) return;
    uint32_t dst;
    if (dead)
        dst = meta[0] + tile_dead[blockIdx.x] + wave_offset(s_dead, w) + (uint32_t)__popcll(dm & lanes_below(lane));
    else
        dst = start[k] + rank[i];
    if (dst >= n) return;   // unreachable while the counts start at zero; never write past the arrays
    kout[dst] = k;
    vout[dst] = i;
    so[dst] = orig_old[i];
}

// cellStart / cellEnd (inclusive) of every non-empty cell, UGNSCudaImpl.cu:144-166.
__global__ void k_cell_ranges(const uint32_t *__restrict__ key, uint32_t n, uint32_t *__restrict__ start,
                              uint32_t *__restrict__ end) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t k = key[i];
    if (i == 0 || key[i - 1] != k) start[k] = i;
    if (i == n - 1 || key[i + 1] != k) end[k] = i;
}

__global__ void k_gather_u32(const uint32_t *__restrict__ src, const uint32_t *__restrict__ idx, uint32_t n,
                             uint32_t *__restrict__ dst) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

// Within one cell the sort leaves particles in previous-slot order (radix) or
// in atomic order (counting);
// rank them by attach id so the final order is (cellId, attachId), the order
// thrust's stable sort of attach-ordered input produces.  Off-grid particles
// share the dead key and keep the stable order.
__global__ void k_cell_fixup(const uint32_t *__restrict__ key, const uint32_t *__restrict__ val,
                             const uint32_t *__restrict__ so, const uint32_t *__restrict__ start,
                             const uint32_t *__restrict__ end, uint32_t n, uint32_t dead_key,
                             uint32_t *__restrict__ perm, uint32_t *__restrict__ orig_new) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t k = key[i];
    uint32_t my = so[i];
    uint32_t dst = i;
    if (k != dead_key) {
        uint32_t s = start[k], e = end[k];
        uint32_t rank = 0;
        for (uint32_t j = s; j <= e; ++j) rank += so[j] < my ? 1u : 0u;
        dst = s + rank;
    }
    perm[dst] = val[i];
    orig_new[dst] = my;
}

// Reference traversal of one particle (UGNSCudaImpl.cu:168-203): the 27
// cells in the reference's offset order (x outermost, z innermost,
// UniformGridNeighborSearch.cpp:33-60), every other particle of a cell in
// slot order that passes 0 < |x_i - x_j| <= h, capped at `cap` entries.
// Nothing after the cap-th hit can change the list, so the walk stops once
// the cap is reached (checked between batches of candidates; hits past the
// cap inside a batch are counted but not stored) and returns min(hits, cap).
template <typename Put>
__device__ uint32_t reference_scan(const float4 *__restrict__ x, const uint32_t *__restrict__ start,
                                   const uint32_t *__restrict__ end, const Grid &g, float3 p, int cx, int cy,
                                   int cz, uint32_t sbase, uint32_t cap, const uint32_t *__restrict__ bases,
                                   uint32_t stride, uint32_t i, Put &&put) {
    const uint32_t gxy = (uint32_t)(g.gx * g.gy);
    uint32_t c = 0;
    // j == i gives d2 == 0, which the test rejects (the reference skips it)
    auto test = [&](uint32_t j, const float4 &xj, uint32_t ztag) {
        if (is_neighbor(dist2_exact(p, xyz(xj)), g.nbr_d2_max)) {
            if (c < cap) put(c, j, xj, ztag);
            ++c;
        }
    };
    constexpr uint32_t B = VTP_REF_BATCH;
    for (int ox = -1; ox <= 1; ++ox) {
        const int X = cx + ox;
        if (X < 0 || X >= g.gx) continue;
        for (int oy = -1; oy <= 1; ++oy) {
            const int Y = cy + oy;
            if (Y < 0 || Y >= g.gy) continue;
            for (int oz = -1; oz <= 1; ++oz) {
                const int Z = cz + oz;
                if (Z < 0 || Z >= g.gz) continue;
                const uint32_t cid = sbase + (uint32_t)Z * gxy + (uint32_t)Y * (uint32_t)g.gx + (uint32_t)X;
                uint32_t j = ldg(start, cid);
                if (j == kEmpty) continue;
                const uint32_t e = ldg(end, cid);
                // the entry tag of this cell's layer: its base minus (layer << 14)
                const uint32_t ztag = bases[(size_t)(oz + 1) * stride + i] - ((uint32_t)(oz + 1) << kNbrOffBits);
                // batches of B; the last one loads slot e again in place of
                // slots past the cell and skips their tests
                for (; j <= e && c < cap; j += B) {
                    float4 a[B];
#pragma unroll
                    for (uint32_t b = 0; b < B; ++b) a[b] = ldg(x, min(j + b, e));
#pragma unroll
                    for (uint32_t b = 0; b < B; ++b)
                        if (j + b <= e) test(j + b, a[b], ztag);
                }
                if (c >= cap) return cap;
            }
        }
    }
    return c;
}

// Neighbour lists.  Cells (X-1, X, X+1) of one (Y, Z) row have consecutive
// keys, so their particles occupy one contiguous slot range.  The reference
// visits the 27 cells with the x offset outermost and keeps the first `cap`
// neighbours it finds.  Two walks give exactly that list:
//   row walk (default): the 9 row ranges in key order, so every list is
//     sorted by slot.  When nothing is cut by the cap this is the
//     reference's set; a particle with more than `cap` neighbours is rebuilt
//     with the reference traversal.  Sorted lists keep the lanes of a wave on
//     neighbouring cache lines in every later neighbour pass (measured at 1M
//     particles: IMM-CT ticks 20 % and DFSPH 12 % faster than with the x+1
//     cells moved to the end of each list).
//   reference walk: the reference traversal for every particle, stopping at
//     the cap: no second walk for the particles at the cap.
// The build counts the particles that reach the cap into kStNbrAtCap;
// k_step_begin moves the count to kStNbrPrev, and the next build takes the
// reference walk when more than 1/kRefOrderShare of the particles reached
// the cap (the choice is read from device memory, so a captured graph adapts
// as well).  Both walks give the reference's sets; the order inside a list
// differs, which changes FP32 rounding in the neighbour sums, so the count is
// part of the carried state.
//
// kLam = 1 (2: mass-weighted) also accumulates the density and lambda sums
// of a PBF tick's first constraint iteration over each list as it is
// written (the walk holds every listed neighbour's position and mass), with
// the lambda pass's own per-neighbour code and order, and stores the pass's
// outputs, static particles' included.
//
// Entries are written as 16-bit (cell layer, offset) pairs from the
// particle's per-layer base slots (layout in vtp_device.h).  A 64-slot group
// in which any offset does not fit 14 bits is marked wide, and the
// kWide instance of the kernel, launched after every build, walks the lists
// of the marked groups again (same walk, same order) and stores them as full
// 32-bit entries; every other group returns at once.  The kWide instance
// neither counts caps nor computes lambda: the first instance already did,
// over the same neighbours in the same order.
template <bool kStaged, int kLam, bool kWide>
__global__ __launch_bounds__(kNbrBuildThreads, VTP_BUILD_MIN_WAVES) void k_build_neighbors(const float4 *__restrict__ x,
                                                                      const uint32_t *__restrict__ key,
                                                                      const uint32_t *__restrict__ start,
                                                                      const uint32_t *__restrict__ end, Grid g,
                                                                      uint32_t n, uint32_t stride, uint32_t cap,
                                                                      uint32_t *__restrict__ nbr,
                                                                      uint32_t *__restrict__ cnt,
                                                                      uint32_t *__restrict__ status,
                                                                      uint32_t walk,
                                                                      const int32_t *__restrict__ fluid_only,
                                                                      PbfLambdaOut lo) {
    extern __shared__ uint16_t s_list[];
    // a thread's staged entries pair up in its own LDS words (entries 2k and
    // 2k + 1 in word k of its column), so no two lanes write one word
    auto stage_slot = [&](uint32_t c) { return ((c >> 1) * blockDim.x + threadIdx.x) * 2u + (c & 1u); };
    const uint32_t i = slot_index();
    uint32_t *const wide_flag = nbr + 3 * (size_t)stride;
    uint4 *const chunks = nbr_chunks_of(nbr, stride);
    uint16_t *const narrow = reinterpret_cast<uint16_t *>(chunks);
    if (kWide && !__builtin_amdgcn_readfirstlane(i < stride ? wide_flag[i >> 6] : 0u)) return;
    // uniform over the grid; walk = 1 / 2 forces the row / reference walk
    const bool ref_walk = walk == 2 || (walk == 0 && status[kStNbrPrev] > n / kRefOrderShare);
    const uint32_t stage = (kStaged && !kWide) ? min(cap, kLam ? kNbrStageLambda : kNbrStage) : 0u;
    float4 xi{};
    float rho = 0, gsx = 0, gsy = 0, gsz = 0, s2 = 0;
    bool sums = false;   // this particle's lambda sums are accumulated
    // A layer's offsets fit 14 bits when its rows span fewer than 2^14
    // slots past its base; checked once per row, a group with any larger
    // span is stored wide.
    bool over = false;
    // ztag = the entry's layer base - (layer << 14): the entry is j - ztag
    // (its top two bits the layer) whenever the offset fits
    auto put = [&](uint32_t c, uint32_t j, const float4 &xj, uint32_t ztag) {
        if (kWide) {
            reinterpret_cast<uint32_t *>(chunks)[nbr_wide_pos(stride, i, c)] = j;
            return;
        }
        const uint16_t e = (uint16_t)(j - ztag);
        if (c < stage) s_list[stage_slot(c)] = e;
        else narrow[nbr_narrow_pos(stride, i, c)] = e;
        if (kLam && sums) pbf_lambda_pair<kLam == 2>(xi, xj, lo.q.k, rho, gsx, gsy, gsz, s2);
    };
    auto restart = [&]() {   // the self term first, as the lambda pass
        rho = xi.w * lo.q.k.w0;
        gsx = gsy = gsz = s2 = 0.0f;
    };
    uint32_t c = 0;
    bool full = false;
    if (i < n) {
        const uint32_t k = key[i];
        bool fluid = false;
        if (kLam) {
            xi = x[i];
            fluid = lo.mat[i] == VTP_MAT_FLUID;
        }
        if (k >= g.dead_key) {
            cnt[i] = kDeadBit;
            if (kLam && fluid)   // no neighbours and no self term (count word's dead bit)
                pbf_lambda_store(i, xi, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, lo.q, lo.lam, lo.dens, lo.errc, lo.sg, lo.xl);
        } else if (fluid_only && fluid_only[i] != VTP_MAT_FLUID) {
            cnt[i] = 0;   // a static particle whose list no kernel of this tick reads
        } else {
            const uint32_t gxy = (uint32_t)(g.gx * g.gy);
            const uint32_t scene = k / g.cells;
            const uint32_t local = k - scene * g.cells;
            const int cz = (int)(local / gxy);
            const uint32_t rem = local - (uint32_t)cz * gxy;
            const int cy = (int)(rem / (uint32_t)g.gx);
            const int cx = (int)(rem - (uint32_t)cy * (uint32_t)g.gx);
            const uint32_t sbase = scene * g.cells;
            const float3 p = kLam ? xyz(xi) : xyz(x[i]);
            sums = kLam && fluid;
            if (sums) restart();
            auto test = [&](uint32_t j, const float4 &xj, uint32_t ztag) {
                // j == i gives d2 == 0, which the test rejects
                if (is_neighbor(dist2_exact(p, xyz(xj)), g.nbr_d2_max)) {
                    if (c < cap) put(c, j, xj, ztag);
                    ++c;
                }
            };
            auto walk = [&](uint32_t j, uint32_t e, uint32_t ztag) {
                constexpr uint32_t B = kLam ? VTP_BUILD_BATCH_LAMBDA : VTP_BUILD_BATCH;   // candidates in flight
                for (; j <= e; j += B) {
                    float4 a[B];
#pragma unroll
                    for (uint32_t b = 0; b < B; ++b) a[b] = ldg(x, min(j + b, e));
#pragma unroll
                    for (uint32_t b = 0; b < B; ++b)
                        if (j + b <= e) test(j + b, a[b], ztag);
                }
            };
            const int X0 = max(cx - 1, 0), X1 = min(cx + 1, g.gx - 1);
            // cell-table entries (start, end) of the 3 cells of row
            // r = 3 (dz + 1) + (dy + 1), rows in key order; cells past X1
            // repeat X1, and out-of-grid rows read as empty
            uint32_t ts[3], te[3];
            auto request = [&](int r) {
                const int Z = cz + r / 3 - 1, Y = cy + r % 3 - 1;
                const bool ok = Z >= 0 && Z < g.gz && Y >= 0 && Y < g.gy;
                // clamped, so that the address is a cell of this scene even where !ok
                const uint32_t row = sbase + (uint32_t)min(max(Z, 0), g.gz - 1) * gxy +
                                     (uint32_t)min(max(Y, 0), g.gy - 1) * (uint32_t)g.gx;
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const uint32_t cell = row + (uint32_t)min(X0 + q, X1);
                    ts[q] = ok ? ldg(start, cell) : kEmpty;
                    te[q] = ok ? ldg(end, cell) : kEmpty;
                }
            };
            // the row's slot range [s, e] from the requested entries (s = kEmpty: empty row)
            auto range = [&](uint32_t &s, uint32_t &e) {
                s = kEmpty;
                e = 0;
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    if (ts[q] != kEmpty) {
                        if (s == kEmpty) s = ts[q];
                        e = te[q];
                    }
                }
            };
            // Each cell layer's base (its lowest slot, the start of its first
            // non-empty row in key order; kEmpty for an empty layer) goes to
            // the list buffer as soon as the layer's rows are known: only the
            // current layer's base stays in a register, and the reference
            // traversal reads them back.
            if (ref_walk) {
                uint32_t zbase = kEmpty;
#pragma unroll 1
                for (int r = 0; r < 9; ++r) {
                    uint32_t s, e;
                    request(r);
                    range(s, e);
                    if (r % 3 == 0) zbase = kEmpty;
                    if (zbase == kEmpty) zbase = s;
                    over |= s != kEmpty && e - zbase > kNbrOffMask;
                    if (r % 3 == 2) nbr[(size_t)(r / 3) * stride + i] = zbase;
                }
                c = reference_scan(x, start, end, g, p, cx, cy, cz, sbase, cap, nbr, stride, i, put);
            } else {
                // The 9 row ranges in key order (z outer, y inner): a sorted
                // list.  The cell-table entries of row r + 1 are requested
                // before row r is walked, and a row's starts and ends are
                // loaded together, so the table loads cost no round trip of
                // their own.
                request(0);
                uint32_t zbase = kEmpty;
#pragma unroll 1
                for (int r = 0; r < 9; ++r) {
                    uint32_t s, e;
                    range(s, e);
                    if (r < 8) request(r + 1);
                    if (r % 3 == 0) zbase = kEmpty;
                    if (zbase == kEmpty) zbase = s;   // before the layer's first entry
                    over |= s != kEmpty && e - zbase > kNbrOffMask;
                    if (s != kEmpty && c <= cap)   // past the cap: rebuilt below
                        walk(s, e, zbase - ((uint32_t)(r / 3) << kNbrOffBits));
                    if (r % 3 == 2) nbr[(size_t)(r / 3) * stride + i] = zbase;
                }
                if (c > cap) {
                    if (sums) restart();
                    c = reference_scan(x, start, end, g, p, cx, cy, cz, sbase, cap, nbr, stride, i, put);
                }
            }
            cnt[i] = c;
            full = cap > 0 && c == cap;
            if (sums) pbf_lambda_store(i, xi, rho, gsx, gsy, gsz, s2, lo.q, lo.lam, lo.dens, lo.errc, lo.sg, lo.xl);
        }
        if (kLam && !fluid) pbf_lambda_store_static(i, xi, lo.lam, lo.dens, lo.errc, lo.sg, lo.xl);
    }
    if (kWide) return;
    // particles that reached the cap steer the walk of the next build
    const uint64_t m = __ballot(full);
    if (m && (threadIdx.x & 63u) == 0) {
        atomicAdd(&status[kStNbrAtCap], (uint32_t)__popcll(m));
        atomicAdd(&status[kStCapSeen], (uint32_t)__popcll(m));
    }
    // one format per 64-slot group (a wave): wide if any offset overflowed
    const bool wide = __ballot(over) != 0;
    if ((threadIdx.x & 63u) == 0 && i < stride) wide_flag[i >> 6] = wide ? 1u : 0u;
    if (kStaged && !wide) {
        // every lane reads back only its own LDS column (no barrier needed)
        // and stores it in whole chunks; stage is a multiple of 8 whenever
        // entries past it went to memory directly, so no chunk mixes the two
        const uint32_t cmax = min(wave_max_u32(c), stage);
        const uint32_t *words = reinterpret_cast<const uint32_t *>(s_list);
        for (uint32_t q = 0; 8 * q < cmax; ++q)
            if (8 * q < c) {
                const uint32_t w = 4 * q * blockDim.x + threadIdx.x;
                chunks[(size_t)q * stride + i] =
                    make_uint4(words[w], words[w + blockDim.x], words[w + 2 * blockDim.x], words[w + 3 * blockDim.x]);
            }
    }
}

// Reference-layout export (UniformGridNeighborSearch.hpp:28-63): attach ids,
// self first, UINT32_MAX padding, neighbours in the reference traversal
// order (offset t of the 27-cell table, then slot).
__global__ void k_export_neighbors(const uint32_t *__restrict__ nbr, const uint32_t *__restrict__ cnt,
                                   const uint32_t *__restrict__ orig, const uint32_t *__restrict__ key, Grid g,
                                   uint32_t n, uint32_t stride, uint32_t M, uint32_t *__restrict__ out_lists,
                                   uint32_t *__restrict__ out_num) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t p = orig[i];
    uint32_t w = cnt[i];
    uint32_t *row = out_lists + (size_t)p * M;
    if (w & kDeadBit) {
        out_num[p] = 0;
        for (uint32_t t = 0; t < M; ++t) row[t] = kEmpty;
        return;
    }
    const uint32_t c = w & ~kDeadBit;
    const uint32_t gxy = (uint32_t)(g.gx * g.gy);
    auto coords = [&](uint32_t kk, int &X, int &Y, int &Z) {
        uint32_t local = kk % g.cells;
        Z = (int)(local / gxy);
        uint32_t rem = local - (uint32_t)Z * gxy;
        Y = (int)(rem / (uint32_t)g.gx);
        X = (int)(rem - (uint32_t)Y * (uint32_t)g.gx);
    };
    int xi, yi, zi;
    coords(key[i], xi, yi, zi);
    // order key: offset index t = (ox+1)*9 + (oy+1)*3 + (oz+1), then slot; insertion sort
    auto okey = [&](uint32_t j) -> uint64_t {
        int X, Y, Z;
        coords(key[j], X, Y, Z);
        uint32_t t = (uint32_t)((X - xi + 1) * 9 + (Y - yi + 1) * 3 + (Z - zi + 1));
        return ((uint64_t)t << 32) | j;
    };
    row[0] = p;
    for (uint32_t a = 0; a < c; ++a) {
        uint32_t ja = nbr_entry(nbr, stride, i, a);
        uint64_t ka = okey(ja);
        uint32_t b = a;
        // row[1..a] holds slot ids sorted so far (translated below)
        while (b > 0 && okey(row[b]) > ka) {
            row[b + 1] = row[b];
            --b;
        }
        row[b + 1] = ja;
    }
    for (uint32_t t = 1; t <= c; ++t) row[t] = orig[row[t]];
    for (uint32_t t = c + 1; t < M; ++t) row[t] = kEmpty;
    out_num[p] = c + 1;
}

// ---------------------------------------------------------------- host side
static inline uint32_t blocks_for(uint32_t n, uint32_t b) { return (n + b - 1) / b; }

void Ugns::allocate(uint32_t n_, const Grid &g_, uint32_t max_neighbors) {
    n = n_;
    g = g_;
    M = max_neighbors;
    cap = M > 0 ? M - 1 : 0;
    stride = (n + 63u) & ~63u;
    uint32_t maxkey = g.dead_key;
    key_bits = 1;
    while (key_bits < 32 && (maxkey >> key_bits) != 0) ++key_bits;
    passes = (key_bits + 7) / 8;
    tiles = blocks_for(n, kRsTile);
    for (int b = 0; b < 2; ++b) {
        key[b].alloc(n);
        val[b].alloc(n);
    }
    spine.alloc((size_t)256 * tiles);
    dtot.alloc(256);
    // counting sort: per-cell counts (zero between searches), ranks, block
    // and tile sums
    count.alloc((size_t)g.dead_key);
    VTP_HIP(hipMemset(count.p, 0, count.bytes()));
    rank.alloc(n);
    bsum.alloc(blocks_for(g.dead_key, kScanBlock));
    tile_dead.alloc(blocks_for(n, kCsThreads));
    meta.alloc(4);
    start.alloc((size_t)g.dead_key + 1);
    end.alloc((size_t)g.dead_key + 1);
    so.alloc(n);
    perm.alloc(n);
    orig_new.alloc(n);
    nbr.alloc(nbr_words(cap, stride));
    cnt.alloc(n);
}

// Input: key[0], val[0] filled by the solver's hash kernel.
// Output: sorted keys in sorted_key(), permutation perm (old slot of each new
// slot) and orig_new (attach id of each new slot).
void Ugns::sort_and_rank(const uint32_t *orig_old, hipStream_t s, StageTimer &tm) {
    if (!radix) {
        const uint32_t ncell = g.dead_key;   // live keys; the dead key's range is set by the spine
        const uint32_t ntile = blocks_for(n, kCsThreads), nblk = blocks_for(ncell, kScanBlock);
        tm.begin(s, "ugns_sort");
        hipLaunchKernelGGL(k_cell_count, dim3(ntile), dim3(kCsThreads), 0, s, key[0].p, n, g.dead_key, count.p,
                           rank.p, tile_dead.p);
        hipLaunchKernelGGL(k_cell_scan_reduce, dim3(nblk), dim3(kCsThreads), 0, s, count.p, ncell, bsum.p);
        hipLaunchKernelGGL(k_cell_scan_spine, dim3(1), dim3(kSpineThreads), 0, s, bsum.p, nblk, tile_dead.p, ntile, n,
                           g.dead_key, start.p, end.p, meta.p);
        hipLaunchKernelGGL(k_cell_scan_ranges, dim3(nblk), dim3(kCsThreads), 0, s, count.p, ncell, bsum.p, start.p,
                           end.p);
        hipLaunchKernelGGL(k_cell_scatter, dim3(ntile), dim3(kCsThreads), 0, s, key[0].p, n, g.dead_key, rank.p,
                           start.p, tile_dead.p, meta.p, orig_old, key[1].p, val[1].p, so.p);
        sorted = 1;
        tm.end(s);
        tm.begin(s, "ugns_ranges");
        hipLaunchKernelGGL(k_cell_fixup, dim3(blocks_for(n, 256)), dim3(256), 0, s, key[1].p, val[1].p, so.p,
                           start.p, end.p, n, g.dead_key, perm.p, orig_new.p);
        tm.end(s);
        return;
    }
    tm.begin(s, "ugns_sort");
    int b = 0;
    for (uint32_t p = 0; p < passes; ++p) {
        int shift = (int)(8 * p);
        hipLaunchKernelGGL(k_radix_hist, dim3(tiles), dim3(kRsThreads), 0, s, key[b].p, n, shift, spine.p, tiles);
        hipLaunchKernelGGL(k_radix_spine, dim3(256), dim3(kRsThreads), 0, s, spine.p, tiles, dtot.p);
        hipLaunchKernelGGL(k_radix_scatter, dim3(tiles), dim3(kRsThreads), 0, s, key[b].p, val[b].p,
                           key[b ^ 1].p, val[b ^ 1].p, n, shift, spine.p, dtot.p, tiles);
        b ^= 1;
    }
    sorted = b;
    tm.end(s);
    tm.begin(s, "ugns_ranges");
    const uint32_t ncell = g.dead_key + 1;
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(ncell, 256)), dim3(256), 0, s, start.p, ncell, kEmpty);
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(ncell, 256)), dim3(256), 0, s, end.p, ncell, kEmpty);
    hipLaunchKernelGGL(k_cell_ranges, dim3(blocks_for(n, 256)), dim3(256), 0, s, key[b].p, n, start.p, end.p);
    hipLaunchKernelGGL(k_gather_u32, dim3(blocks_for(n, 256)), dim3(256), 0, s, orig_old, val[b].p, n, so.p);
    hipLaunchKernelGGL(k_cell_fixup, dim3(blocks_for(n, 256)), dim3(256), 0, s, key[b].p, val[b].p, so.p,
                       start.p, end.p, n, g.dead_key, perm.p, orig_new.p);
    tm.end(s);
}

void Ugns::build_neighbors(const float4 *x, hipStream_t s, StageTimer &tm, const int32_t *fluid_only,
                           const PbfLambdaOut *lam) {
    if (!status) throw std::runtime_error("neighbour build without status words");
    static_rows = fluid_only == nullptr;
    tm.begin(s, "ugns_neighbors");
    // the stage is read back in whole 8-entry chunks: a stage cut by a small
    // cap is rounded up to one
    const uint32_t staged = std::min(cap, lam ? kNbrStageLambda : kNbrStage);
    const size_t lds = (size_t)((staged + 7) / 8) * 8 * kNbrBuildThreads * sizeof(uint16_t);
    const dim3 grid(blocks_for(n, kNbrBuildThreads)), block(kNbrBuildThreads);
#define VTP_BUILD(LAM, OUT)                                                                                   \
    hipLaunchKernelGGL((k_build_neighbors<true, LAM, false>), grid, block, lds, s, x, key[sorted].p, start.p,   \
                       end.p, g, n, stride, cap, nbr.p, cnt.p, status, walk, fluid_only, OUT)
    if (!lam) VTP_BUILD(0, PbfLambdaOut{});
    else if (lam->q.mass_weighted) VTP_BUILD(2, *lam);
    else VTP_BUILD(1, *lam);
#undef VTP_BUILD
    // groups whose offsets overflowed 16 bits: the same lists as 32-bit entries
    hipLaunchKernelGGL((k_build_neighbors<false, 0, true>), grid, block, 0, s, x, key[sorted].p, start.p, end.p, g,
                       n, stride, cap, nbr.p, cnt.p, status, walk, fluid_only, PbfLambdaOut{});
    tm.end(s);
}

void Ugns::export_reference(const uint32_t *orig, uint32_t *d_lists, uint32_t *d_num, hipStream_t s) {
    hipLaunchKernelGGL(k_export_neighbors, dim3(blocks_for(n, 256)), dim3(256), 0, s, nbr.p, cnt.p, orig,
                       key[sorted].p, g, n, stride, M, d_lists, d_num);
}

void launch_fill_u32(uint32_t *a, uint32_t n, uint32_t v, hipStream_t s) {
    if (n) hipLaunchKernelGGL(k_fill_u32, dim3(blocks_for(n, 256)), dim3(256), 0, s, a, n, v);
}

}  // namespace vtp
