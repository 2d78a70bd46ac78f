// state.hip - binary state image of a solver context (checkpoint / resume).
//
// The reference has no resume path: its exporter writes positions and colour
// only (ExportUtil.cpp:28-88).  The image holds exactly the state one tick
// carries into the next, in the current slot order together with the slot ->
// attach map, so a context that loads it ticks on bit for bit:
//   PBF / DFSPH   orig, X[0] (w = mass), V[0] (w = volume)
//   IMM           + frac[0], uphase[0], rest2[0], color4[0]   (the arrays the
//                   IMM tick gathers into the new slot order)
//   IMM-CT        + Q[0]
// Everything else a tick reads is rebuilt at its start (hash, sort, lists,
// per-slot material/scene/object tables) or derived from the static
// attach-order tables, except the count of particles that reached the
// neighbour cap in the last build: it selects the walk of the next build,
// which fixes the order inside the lists, so it travels in the header.
// Phase-major arrays are stored packed ([k][n]) rather than with the device
// stride.
#include <cstring>
#include <vector>

#include "vtp_internal.h"

using namespace vtp;

namespace {

constexpr char kMagic[8] = {'V', 'T', 'P', 'S', 'T', 'A', 'T', 'E'};
constexpr uint32_t kVersion = 1;

// Per-slot material / scene / object from the attach id of each slot.
__global__ void k_slot_tables(const uint32_t *__restrict__ orig, const int32_t *__restrict__ mat_attach,
                              const uint32_t *__restrict__ scene_attach, const uint32_t *__restrict__ obj_attach,
                              uint32_t n, int32_t *__restrict__ mat, uint32_t *__restrict__ scene,
                              uint32_t *__restrict__ obj_of) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t a = orig[i];
    mat[i] = mat_attach[a];
    scene[i] = scene_attach[a];
    if (obj_of) obj_of[i] = obj_attach[a];
}

uint64_t fnv1a(uint64_t h, const void *p, size_t len) {
    const unsigned char *b = static_cast<const unsigned char *>(p);
    for (size_t i = 0; i < len; ++i) {
        h ^= b[i];
        h *= 1099511628211ull;
    }
    return h;
}

uint64_t layout_hash(const Ctx &c) {
    uint64_t h = 1469598103934665603ull;
    const int32_t solver = c.p.solver;
    h = fnv1a(h, &solver, 4);
    h = fnv1a(h, &c.n, 4);
    h = fnv1a(h, &c.P, 4);
    for (const Object &o : c.objects) {
        const uint32_t w[4] = {o.start, o.end, o.scene, (uint32_t)o.mat};
        h = fnv1a(h, w, sizeof w);
    }
    return h;
}

bool is_imm(const Ctx &c) { return c.p.solver == VTP_IMM || c.p.solver == VTP_IMMCT; }

// One section of the image: `rows` device rows of `row_bytes`, row r at
// dev + r * pitch_bytes.
struct Section {
    void *dev;
    size_t row_bytes, pitch_bytes, rows;
    size_t bytes() const { return row_bytes * rows; }
};

std::vector<Section> sections(Ctx &c) {
    const size_t n = c.n, st = c.ugns.stride, P = c.P;
    std::vector<Section> s;
    s.push_back({c.orig.p, n * 4, n * 4, 1});
    s.push_back({c.X[0].p, n * 16, n * 16, 1});
    s.push_back({c.V[0].p, n * 16, n * 16, 1});
    if (is_imm(c)) {
        s.push_back({c.frac[0].p, n * 4, st * 4, P});
        s.push_back({c.uphase[0].p, n * 16, st * 16, P});
        s.push_back({c.rest2[0].p, n * 4, n * 4, 1});
        s.push_back({c.color4[0].p, n * 16, n * 16, 1});
        if (c.p.solver == VTP_IMMCT) s.push_back({c.Q[0].p, n * 4, st * 4, 9});
    }
    return s;
}

// Rows are copied one 1D hipMemcpyAsync each (a section has at most 9), which
// behaves the same under every runtime interposer; a 2D copy into pageable
// host memory came back as zeros under rocprofv3's kernel tracing.
void copy_rows_d2h(unsigned char *host, const Section &s, hipStream_t st) {
    for (size_t r = 0; r < s.rows; ++r)
        VTP_HIP(hipMemcpyAsync(host + r * s.row_bytes, static_cast<const unsigned char *>(s.dev) + r * s.pitch_bytes,
                               s.row_bytes, hipMemcpyDeviceToHost, st));
}
void copy_rows_h2d(const unsigned char *host, const Section &s, hipStream_t st) {
    for (size_t r = 0; r < s.rows; ++r)
        VTP_HIP(hipMemcpyAsync(static_cast<unsigned char *>(s.dev) + r * s.pitch_bytes, host + r * s.row_bytes,
                               s.row_bytes, hipMemcpyHostToDevice, st));
}

size_t payload_bytes(Ctx &c) {
    size_t b = 0;
    for (const Section &s : sections(c)) b += s.bytes();
    return b;
}

vtp_status fail(Ctx &c, vtp_status s, const std::string &msg) {
    c.err = msg;
    return s;
}

template <typename F>
vtp_status guarded(Ctx &c, F &&f) {
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

}  // namespace

extern "C" vtp_status vtp_state_size(vtp_ctx *ctx, size_t *bytes) {
    if (!ctx) return VTP_EINVAL;
    Ctx &c = ctx->c;
    if (!bytes) return fail(c, VTP_EINVAL, "vtp_state_size: null output");
    if (!c.initialized) return fail(c, VTP_ESTATE, "vtp_state_size: not initialized");
    *bytes = sizeof(vtp_state_header) + payload_bytes(c);
    return VTP_OK;
}

namespace vtp {
size_t state_image_bytes(Ctx &c) { return sizeof(vtp_state_header) + payload_bytes(c); }
vtp_status save_state_image(Ctx &c, void *host_dst, size_t bytes);
vtp_status load_state_image(Ctx &c, const void *host_src, size_t bytes);
}  // namespace vtp

extern "C" vtp_status vtp_save_state(vtp_ctx *ctx, void *host_dst, size_t bytes) {
    if (!ctx) return VTP_EINVAL;
    return save_state_image(ctx->c, host_dst, bytes);
}

extern "C" vtp_status vtp_load_state(vtp_ctx *ctx, const void *host_src, size_t bytes) {
    if (!ctx) return VTP_EINVAL;
    return load_state_image(ctx->c, host_src, bytes);
}

vtp_status vtp::save_state_image(Ctx &c, void *host_dst, size_t bytes) {
    if (!c.initialized) return fail(c, VTP_ESTATE, "vtp_save_state: not initialized");
    if (!host_dst) return fail(c, VTP_EINVAL, "vtp_save_state: null destination");
    const size_t need = sizeof(vtp_state_header) + payload_bytes(c);
    if (bytes < need)
        return fail(c, VTP_EINVAL, "vtp_save_state: buffer too small (need " + std::to_string(need) + " bytes)");
    return guarded(c, [&]() -> vtp_status {
        read_status(c);
        vtp_state_header h;
        std::memset(&h, 0, sizeof h);
        std::memcpy(h.magic, kMagic, 8);
        h.version = kVersion;
        h.solver = c.p.solver;
        h.n = c.n;
        h.phases = c.P;
        h.objects = (uint32_t)c.objects.size();
        h.crash_flags = c.h_status[kStCrash];
        h.layout_hash = layout_hash(c);
        h.steps = c.steps;
        h.sim_time = c.sim_time;
        h.total_div_iterations = c.total_div;
        h.total_incomp_iterations = c.total_inc;
        h.iter_hint_div = c.iter_hint_div;
        h.iter_hint_incomp = c.iter_hint_inc;
        h.nbr_at_cap = c.h_status[kStNbrAtCap];
        h.payload_bytes = need - sizeof h;
        unsigned char *out = static_cast<unsigned char *>(host_dst);
        std::memcpy(out, &h, sizeof h);
        size_t off = sizeof h;
        for (const Section &s : sections(c)) {
            copy_rows_d2h(out + off, s, c.stream);
            off += s.bytes();
        }
        VTP_HIP(hipStreamSynchronize(c.stream));
        return VTP_OK;
    });
}

vtp_status vtp::load_state_image(Ctx &c, const void *host_src, size_t bytes) {
    if (!c.initialized) return fail(c, VTP_ESTATE, "vtp_load_state: not initialized");
    if (!host_src || bytes < sizeof(vtp_state_header)) return fail(c, VTP_EINVAL, "vtp_load_state: no header");
    vtp_state_header h;
    std::memcpy(&h, host_src, sizeof h);
    if (std::memcmp(h.magic, kMagic, 8) != 0) return fail(c, VTP_EINVAL, "vtp_load_state: not a state image");
    if (h.version != kVersion)
        return fail(c, VTP_EINVAL, "vtp_load_state: unsupported version " + std::to_string(h.version));
    if (h.solver != c.p.solver) return fail(c, VTP_EINVAL, "vtp_load_state: image is of another solver type");
    if (h.n != c.n || h.phases != c.P || h.objects != c.objects.size() || h.layout_hash != layout_hash(c))
        return fail(c, VTP_EINVAL, "vtp_load_state: attached objects differ from the image's");
    const size_t payload = payload_bytes(c);
    if (h.payload_bytes != payload || bytes != sizeof h + payload)
        return fail(c, VTP_EINVAL, "vtp_load_state: size mismatch (expected " + std::to_string(sizeof h + payload) +
                                       " bytes)");
    const unsigned char *in = static_cast<const unsigned char *>(host_src);
    // the slot -> attach map must be a permutation: kernels index attach tables with it
    {
        const uint32_t n = c.n;
        std::vector<uint8_t> seen(n, 0);
        const unsigned char *po = in + sizeof h;
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t a;
            std::memcpy(&a, po + 4 * (size_t)i, 4);
            if (a >= n || seen[a]) return fail(c, VTP_EINVAL, "vtp_load_state: slot map is not a permutation");
            seen[a] = 1;
        }
    }
    return guarded(c, [&]() -> vtp_status {
        VTP_HIP(hipStreamSynchronize(c.stream));
        size_t off = sizeof h;
        for (const Section &s : sections(c)) {
            copy_rows_h2d(in + off, s, c.stream);
            off += s.bytes();
        }
        hipLaunchKernelGGL(k_slot_tables, dim3(div_up(c.n, 256)), dim3(256), 0, c.stream, c.orig.p, c.mat_attach.p,
                           c.scene_attach.p, c.obj_attach.p, c.n, c.mat.p, c.scene.p,
                           c.obj_attach.p ? c.obj_of.p : nullptr);
        VTP_HIP(hipGetLastError());
        // sticky crash bits travel with the state; the per-tick words restart
        uint32_t words[kStWords] = {0};
        words[kStCrash] = h.crash_flags;
        words[kStNbrAtCap] = h.nbr_at_cap;
        VTP_HIP(hipMemcpyAsync(c.status.p, words, sizeof words, hipMemcpyHostToDevice, c.stream));
        VTP_HIP(hipStreamSynchronize(c.stream));
        c.steps = h.steps;
        c.sim_time = (float)h.sim_time;
        c.total_div = h.total_div_iterations;
        c.total_inc = h.total_incomp_iterations;
        c.iter_hint_div = h.iter_hint_div ? h.iter_hint_div : 1;
        c.iter_hint_inc = h.iter_hint_incomp ? h.iter_hint_incomp : 1;
        c.crashed = (h.crash_flags & ~(uint32_t)VTP_CRASH_OUT_OF_GRID) != 0;
        c.lists_stale = true;      // the lists belong to the replaced state until the next tick
        c.adj_state.reset();       // adjoint checkpoints of an earlier rollout no longer apply
        read_status(c);
        return VTP_OK;
    });
}
