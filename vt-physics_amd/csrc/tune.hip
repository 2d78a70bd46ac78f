// tune.hip - launch-shape autotuner for the neighbour passes.
//
// For every candidate workgroup size the tuner runs two ticks from the same
// saved state with hipEvent stage timing, keeps the fastest size per stage,
// and restores the state, so tuning does not advance or perturb the
// simulation.  PBF and DFSPH only (their state is X, V and the slot map).
#include "vtp_internal.h"

namespace vtp {

void autotune(Ctx &c) {
    if (!c.initialized) return;
    if (c.p.solver != VTP_PBF && c.p.solver != VTP_DFSPH) return;
    const uint32_t n = c.n;
    const std::vector<std::string> stages =
        c.p.solver == VTP_PBF
            ? std::vector<std::string>{"pbf_lambda", "pbf_dx", "pbf_xsph"}
            : std::vector<std::string>{"df_precompute", "df_delta_kappa", "df_update_vadv", "df_vis_normal"};
    // timer names of DFSPH loop stages carry the solve prefix
    auto timed_name = [&](const std::string &s) -> std::vector<std::string> {
        if (s == "df_delta_kappa") return {"df_div_delta_kappa", "df_inc_delta_kappa"};
        if (s == "df_update_vadv") return {"df_div_update_vadv", "df_inc_update_vadv"};
        if (s == "df_vis_normal") return {"df_vis_normal", "df_tension_apply"};
        return {s};
    };
    DevBuf<float4> sx, sv;
    DevBuf<uint32_t> so;
    DevBuf<uint32_t> sstat;
    sx.alloc(n);
    sv.alloc(n);
    so.alloc(n);
    sstat.alloc(kStWords);
    VTP_HIP(hipMemcpyAsync(sx.p, c.X[0].p, n * sizeof(float4), hipMemcpyDeviceToDevice, c.stream));
    VTP_HIP(hipMemcpyAsync(sv.p, c.V[0].p, n * sizeof(float4), hipMemcpyDeviceToDevice, c.stream));
    VTP_HIP(hipMemcpyAsync(so.p, c.orig.p, n * 4, hipMemcpyDeviceToDevice, c.stream));
    VTP_HIP(hipMemcpyAsync(sstat.p, c.status.p, kStWords * 4, hipMemcpyDeviceToDevice, c.stream));
    const uint64_t total_div = c.total_div, total_inc = c.total_inc;
    const uint32_t hint_div = c.iter_hint_div, hint_inc = c.iter_hint_inc;
    const bool timing_was = c.timer.enabled;
    const bool graph_was = c.opt_graph;
    c.opt_graph = false;
    std::map<std::string, std::pair<double, uint32_t>> best;
    const uint32_t cand[] = {64, 128, 256};   // the passes are built for at most 256 threads
    for (uint32_t b : cand) {
        for (const auto &s : stages) c.block_of[s] = b;
        c.timer.reset();
        c.timer.enabled = false;
        solver_tick(c);   // warm-up at this shape
        c.timer.enabled = true;
        solver_tick(c);
        c.timer.collect();
        for (const auto &s : stages) {
            double t = 0;
            for (const auto &nm : timed_name(s)) t += c.timer.ms.count(nm) ? c.timer.ms[nm] : 0.0;
            if (!best.count(s) || t < best[s].first) best[s] = {t, b};
        }
        // back to the saved state
        VTP_HIP(hipMemcpyAsync(c.X[0].p, sx.p, n * sizeof(float4), hipMemcpyDeviceToDevice, c.stream));
        VTP_HIP(hipMemcpyAsync(c.V[0].p, sv.p, n * sizeof(float4), hipMemcpyDeviceToDevice, c.stream));
        VTP_HIP(hipMemcpyAsync(c.orig.p, so.p, n * 4, hipMemcpyDeviceToDevice, c.stream));
        VTP_HIP(hipMemcpyAsync(c.status.p, sstat.p, kStWords * 4, hipMemcpyDeviceToDevice, c.stream));
        void launch_mat_from_orig(Ctx &);
        launch_mat_from_orig(c);
    }
    for (const auto &kv : best) c.block_of[kv.first] = kv.second.second;
    c.total_div = total_div;
    c.total_inc = total_inc;
    c.iter_hint_div = hint_div;
    c.iter_hint_inc = hint_inc;
    c.timer.reset();
    c.timer.enabled = timing_was;
    c.opt_graph = graph_was;
    VTP_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace vtp
