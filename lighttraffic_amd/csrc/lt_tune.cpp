// lt_tune.cpp — autotuner of batched PPR (lt_tune_ppr_batch), host code over
// the engine's own C ABI.  No counterpart in the reference, which runs one
// source per program (src/walk/SingleSource.h:6-20) with compile-time knobs.
//
// Every configuration gives the same endpoints (walker ids and Philox draws
// do not depend on it), so the tuner only trades speed against PCIe bytes.
// It times each configuration on every sample_every-th query (a warm second
// job), then the three fastest and the fastest of each copy mode again on a
// sample four times larger, since a small sample walks few steps per
// partition load and so over-weights the copies; then the fastest of that
// second stage with each fixed number of walkers per walk-kernel lane.  The
// pick is the fastest of stages 2 and 3.  Each partition size is tried with
// the reference's split (raw bytes per partition) and, for block copies,
// with the split sized for block slots of that size
// (lt_partition_build_blocks), whose images fill their slots.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "lt_internal.h"

namespace {

struct Sample {
    std::vector<uint32_t> src;
    std::vector<uint64_t> cnt;
};

Sample take_sample(uint32_t nq, const uint32_t *sources, const uint64_t *counts, uint32_t every) {
    Sample s;
    for (uint32_t q = 0; q < nq; q += std::max(1u, every)) {
        s.src.push_back(sources[q]);
        s.cnt.push_back(counts[q]);
    }
    return s;
}

// The partition table of trial t: the reference's split of S = partition_mb,
// or (split 1) the split for block slots of that size.
lt_status build_table(const lt_graph *g, const lt_tune_trial &t, lt_partition **p) {
    return t.split ? lt_partition_build_blocks(g, (uint64_t)t.partition_mb << 20, p)
                   : lt_partition_build(g, (uint64_t)t.partition_mb << 20, p);
}

// Running time of a warm second job of `t`'s configuration on sample `s`.
lt_status time_trial(const lt_graph *g, const lt_partition *p, const Sample &s, float prob, lt_tune_trial &t) {
    lt_config cfg;
    lt_config_default(&cfg);
    cfg.num_device_partitions = t.slots;
    cfg.span = t.span;
    cfg.device_schedule = t.device_schedule;
    cfg.slice_chunks = t.slice_chunks;
    cfg.zc_batch = t.zc_batch;
    cfg.resident_batch = t.zc_batch;  // one batch size for zero-copy and resident launches
    cfg.pack_partitions = t.pack_partitions;
    cfg.walkers_per_thread = t.walkers_per_thread;
    lt_engine *e = nullptr;
    lt_status st = lt_engine_create(g, p, &cfg, &e);
    if (st != LT_OK) return st;
    const uint32_t n = (uint32_t)s.src.size();
    st = lt_run_ppr_batch(e, n, s.src.data(), s.cnt.data(), prob, nullptr);  // fills the cache
    if (st == LT_OK) st = lt_run_ppr_batch(e, n, s.src.data(), s.cnt.data(), prob, nullptr);
    lt_stats stats;
    if (st == LT_OK) st = lt_engine_stats(e, &stats);
    lt_engine_destroy(e);
    if (st != LT_OK) return st;
    t.cache_slots = (uint32_t)stats.cache_slots;
    t.running_ms = stats.running_ms;
    t.h2d_bytes = stats.h2d_bytes_explicit + stats.h2d_bytes_zerocopy;
    t.iterations = stats.iterations;
    return LT_OK;
}

}  // namespace

extern "C" lt_status lt_tune_ppr_batch(const lt_graph *g, uint32_t nq, const uint32_t *sources, const uint64_t *counts,
                                       float prob, uint64_t cache_bytes, const uint32_t *sizes_mb, uint32_t nsizes,
                                       uint32_t sample_every, lt_tune_trial *best_out, lt_tune_trial *trials_out,
                                       uint32_t trials_cap, uint32_t *ntrials_out) {
    if (!g || nq == 0 || !sources || !counts || !best_out || !(prob > 0.0f && prob <= 1.0f) || cache_bytes == 0 ||
        (sizes_mb && nsizes == 0)) {
        lt::set_error("lt_tune_ppr_batch: null argument, no query, no cache, no partition size or prob not in (0, 1]");
        return LT_E_ARG;
    }
    static const uint32_t kDefaultSizesMb[] = {16, 32, 64};
    const std::vector<uint32_t> kSizesMb = sizes_mb ? std::vector<uint32_t>(sizes_mb, sizes_mb + nsizes)
                                                    : std::vector<uint32_t>(kDefaultSizesMb, kDefaultSizesMb + 3);
    static const int32_t kSpans[] = {LT_SPAN_PARTITION, LT_SPAN_RESIDENT};
    static const int32_t kSlices[] = {2048, 8192};
    static const int32_t kZcBatches[] = {1, 16};
    static const int32_t kPacks[] = {0, 1, 2};
    const uint32_t every = std::max(1u, sample_every);
    const Sample s1 = take_sample(nq, sources, counts, every);
    std::vector<lt_tune_trial> trials;
    lt_graph *gw = const_cast<lt_graph *>(g);  // only its cache of packed images changes
    for (uint32_t mb : kSizesMb)
    for (int32_t split : {0, 1}) {
        lt_tune_trial proto{};
        proto.partition_mb = mb;
        proto.split = split;
        lt_partition *p = nullptr;
        lt_status st = build_table(g, proto, &p);
        if (st != LT_OK) return st;
        const uint32_t P = lt_partition_count(p);
        // the same cache bytes: raw slots of the table's partition size
        const uint32_t slots =
            (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(P, cache_bytes / lt_partition_bytes(p)));
        for (int32_t span : kSpans)
            for (int32_t dev : {0, 1})
                for (int32_t sl : kSlices)
                    for (int32_t zb : kZcBatches)
                        for (int32_t pk : kPacks) {
                            if (dev && sl != kSlices[0]) continue;  // the device loop consumes whole queues
                            if (zb != kZcBatches[0] && (dev || span != LT_SPAN_RESIDENT))
                                continue;  // zero-copy batches need host scheduling and the resident span
                            if (dev && pk == 1) continue;  // the device loop has no unpack step
                            if (split && pk != 2) continue;  // the split serves block slots
                            lt_tune_trial t = proto;
                            t.stage = 1;
                            t.slots = slots;
                            t.partitions = P;
                            t.span = span;
                            t.device_schedule = dev;
                            t.slice_chunks = sl;
                            t.zc_batch = zb;
                            t.pack_partitions = pk;
                            if ((st = time_trial(g, p, s1, prob, t)) != LT_OK) {
                                lt_partition_free(p);
                                return st;
                            }
                            trials.push_back(t);
                        }
        lt_graph_drop_packed(gw);  // the next partition size needs images of its own
        lt_partition_free(p);
    }
    // stage 2: the three fastest and the fastest of each copy mode (block
    // copies on either split), on a sample four times larger
    std::vector<lt_tune_trial> top(trials);
    std::sort(top.begin(), top.end(), [](const lt_tune_trial &a, const lt_tune_trial &b) {
        return a.running_ms < b.running_ms;
    });
    std::vector<lt_tune_trial> pick(top.begin(), top.begin() + std::min<size_t>(3, top.size()));
    for (int32_t mode = 0; mode < 4; mode++) {  // raw, gap-coded, block; block on the split for its slots
        const int32_t pk = mode < 3 ? kPacks[mode] : 2, split = mode == 3;
        for (const lt_tune_trial &t : top) {  // sorted, so the first of a mode is its fastest
            if (t.pack_partitions != pk || t.split != split) continue;
            const bool have = std::any_of(pick.begin(), pick.end(), [&](const lt_tune_trial &x) {
                return memcmp(&x, &t, sizeof(t)) == 0;
            });
            if (!have) pick.push_back(t);
            break;
        }
    }
    std::vector<lt_tune_trial> last;
    const Sample s2 = take_sample(nq, sources, counts, std::max(1u, every / 4));
    if (every > 1) {
        for (uint32_t mb : kSizesMb)
        for (int32_t split : {0, 1}) {
            auto same = [&](const lt_tune_trial &t) { return t.partition_mb == mb && t.split == split; };
            if (!std::any_of(pick.begin(), pick.end(), same)) continue;
            lt_partition *p = nullptr;
            lt_status st = build_table(g, *std::find_if(pick.begin(), pick.end(), same), &p);
            if (st != LT_OK) return st;
            for (lt_tune_trial t : pick) {
                if (!same(t)) continue;
                t.stage = 2;
                if ((st = time_trial(g, p, s2, prob, t)) != LT_OK) {
                    lt_partition_free(p);
                    return st;
                }
                last.push_back(t);
            }
            lt_graph_drop_packed(gw);
            lt_partition_free(p);
        }
    } else {
        last = trials;
    }
    lt_tune_trial best = last[0];
    for (const lt_tune_trial &t : last)
        if (t.running_ms < best.running_ms) best = t;
    if (every > 1) trials.insert(trials.end(), last.begin(), last.end());
    // stage 3: the pick with 1, 2 and 4 walkers per lane (the device loop's
    // default is 1, so that variant is the pick itself there)
    {
        lt_partition *p = nullptr;
        lt_status st = build_table(g, best, &p);
        if (st != LT_OK) return st;
        const lt_tune_trial pick = best;
        for (int32_t wpt : {1, 2, 4}) {
            if (pick.device_schedule && wpt == 1) continue;
            lt_tune_trial t = pick;
            t.stage = 3;
            t.walkers_per_thread = wpt;
            if ((st = time_trial(g, p, s2, prob, t)) != LT_OK) {
                lt_partition_free(p);
                return st;
            }
            trials.push_back(t);
            if (t.running_ms < best.running_ms) best = t;
        }
        lt_graph_drop_packed(gw);
        lt_partition_free(p);
    }
    *best_out = best;
    if (trials_out) memcpy(trials_out, trials.data(), sizeof(lt_tune_trial) * std::min<size_t>(trials_cap, trials.size()));
    if (ntrials_out) *ntrials_out = (uint32_t)trials.size();
    return LT_OK;
}
