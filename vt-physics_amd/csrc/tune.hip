// tune.hip - launch-shape autotuner for the neighbour passes of every solver.
//
// Each candidate shape runs two ticks from the same state image, the first
// untimed (warm-up at that shape), the second with the stage hipEvents, and
// the image is loaded back after every candidate: tuning neither advances
// nor perturbs the simulation (positions, velocities, solver fields, time,
// step and iteration counters, crash and cap words are the image's again,
// and the lists are stale as after any load, so nothing the probe ticks
// computed stays readable).
//
// Mode 1: the workgroup size of each neighbour pass, from {64, 128, 256}.
// A particle's sums do not depend on it, so results stay bit-identical.
// Mode 2: also 1, 2 or 4 lanes per particle in the PBF passes (with more
// than one lane the first constraint iteration is not folded into the
// neighbour build) or in the pressure-solve passes of DFSPH, IMM and IMM-CT
// (delta/kappa and the v* update), chosen by the summed time of the build
// and the passes.
// More lanes change the order of a particle's neighbour sums, so results
// then agree with mode 1 within FP32 rounding, not bit for bit.
#include "vtp_internal.h"

namespace vtp {

namespace {

struct Pass {
    const char *block;                // key of Ctx::block_of
    std::vector<const char *> timed;  // stage-timer names of its launches
};

std::vector<Pass> passes_of(const Ctx &c) {
    const Pass dk{"df_delta_kappa", {"df_div_delta_kappa", "df_inc_delta_kappa"}};
    const Pass up{"df_update_vadv", {"df_div_update_vadv", "df_inc_update_vadv"}};
    switch (c.p.solver) {
        case VTP_PBF:
            return {{"pbf_lambda", {"pbf_lambda"}}, {"pbf_dx", {"pbf_dx"}}, {"pbf_xsph", {"pbf_xsph"}}};
        case VTP_DFSPH:
            return {{"df_precompute", {"df_precompute"}}, dk, up, {"df_vis_normal", {"df_vis_normal"}}};
        default: {
            std::vector<Pass> v{dk, up, {"imm_precompute", {"imm_precompute"}},
                                {"imm_vis_normal", {"imm_vis_normal"}},
                                {"imm_tension_phase_vel", {"imm_tension_phase_vel"}},
                                {"imm_fluxes", {"imm_fluxes"}}};
            if (c.p.solver == VTP_IMMCT) {
                v.push_back({"immct_conformation", {"immct_conformation"}});
                v.push_back({"immct_stress_acc", {"immct_stress_acc"}});
            }
            return v;
        }
    }
}

double timed_ms(const StageTimer &t, const std::vector<const char *> &names) {
    double ms = 0;
    for (const char *nm : names) {
        auto it = t.ms.find(nm);
        if (it != t.ms.end()) ms += it->second;
    }
    return ms;
}

}  // namespace

void autotune(Ctx &c, int mode) {
    if (!c.initialized) return;
    const std::vector<Pass> passes = passes_of(c);
    std::vector<unsigned char> image(state_image_bytes(c));
    if (save_state_image(c, image.data(), image.size()) != VTP_OK)
        throw std::runtime_error("autotune: state image: " + c.err);
    const bool timing_was = c.timer.enabled, graph_was = c.opt_graph;
    const uint32_t lanes_was = c.opt_lanes, hint_phase = c.iter_hint_phase;
    const std::map<std::string, uint32_t> blocks_was = c.block_of;
    c.opt_graph = false;
    c.timer.reset();
    const uint32_t cand_block[] = {64, 128, 256};   // the passes are built for at most 256 threads
    std::vector<uint32_t> cand_lanes{1};
    if (mode == 2) cand_lanes = {1, 2, 4};
    // per lane count: the best block of each pass and the summed time
    double best_total = -1;
    uint32_t best_lanes = c.opt_lanes;
    std::map<std::string, uint32_t> best_blocks;
    std::string failure;
    try {
        for (uint32_t L : cand_lanes) {
            c.opt_lanes = L;
            std::map<std::string, std::pair<double, uint32_t>> best;
            double build_ms = 0;
            for (uint32_t b : cand_block) {
                for (const Pass &p : passes) c.block_of[p.block] = b;
                c.timer.reset();
                c.timer.enabled = false;
                solver_tick(c);   // warm-up at this shape
                c.timer.enabled = true;
                solver_tick(c);
                c.timer.collect();
                for (const Pass &p : passes) {
                    const double t = timed_ms(c.timer, p.timed);
                    auto it = best.find(p.block);
                    if (it == best.end() || t < it->second.first) best[p.block] = {t, b};
                }
                build_ms += timed_ms(c.timer, {"ugns_neighbors"}) / 3.0;
                c.timer.enabled = false;
                if (load_state_image(c, image.data(), image.size()) != VTP_OK)
                    throw std::runtime_error("autotune: restoring the state image: " + c.err);
            }
            double total = build_ms;
            for (const auto &kv : best) total += kv.second.first;
            if (best_total < 0 || total < best_total) {
                best_total = total;
                best_lanes = L;
                best_blocks.clear();
                for (const auto &kv : best) best_blocks[kv.first] = kv.second.second;
            }
        }
    } catch (const std::exception &e) {
        failure = e.what();
    }
    c.timer.reset();
    c.timer.enabled = timing_was;
    c.opt_graph = graph_was;
    c.iter_hint_phase = hint_phase;
    if (!failure.empty()) {
        c.block_of = blocks_was;
        c.opt_lanes = lanes_was;
        throw std::runtime_error(failure);
    }
    c.block_of = blocks_was;
    for (const auto &kv : best_blocks) c.block_of[kv.first] = kv.second;
    c.opt_lanes = best_lanes;
    pbf_capture_reset(c);   // a captured tick holds the old launch shapes
    VTP_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace vtp
