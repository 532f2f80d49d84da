// lt_cli.cpp — the pagerank / ppr / genericwalk programs on top of
// liblighttraffic.so.  Built three times with LT_APP_NAME set; stdout lines
// keep the reference formats (src/Config.h:88, src/LightTraffic.h:26-30,
// src/Scheduler.h:440-453, app/pagerank.cu:119-135, app/ppr.cu:96-112) so
// the reference's benchmark greps (benchmark/figure9/run.sh:40,
// benchmark/figure13/iter_analysis.sh:15-17) read them unchanged.
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <string>
#include <vector>

#include "lighttraffic.h"
#include "lt_cli.h"

#ifndef LT_APP_NAME
#define LT_APP_NAME "pagerank"
#endif

static int fail(lt_status s, const char *what) {
    fprintf(stderr, "%s: %s (%s)\n", what, lt_status_str(s), lt_last_error());
    return 1;
}

int main(int argc, char **argv) {
    // Before the HIP runtime starts: at least eight hardware queues (HIP's
    // default is 4), so that a packed copy's DMA stream does not share a queue
    // with the walk stream.  A larger value the caller set is kept.
    const char *hwq = getenv("GPU_MAX_HW_QUEUES");
    if (!hwq || atoi(hwq) < 8) setenv("GPU_MAX_HW_QUEUES", "8", 1);
    const std::string app = LT_APP_NAME;
    CliConfig c;
    if (!cli_parse(argc, argv, c)) {
        fprintf(stderr,
                "usage: %s -f <graph.bcsr> [-m [-s MB]] [-p slots] [-b walkers] [-w walkers] [-l length]\n"
                "          [--runs n] [--source v] [--prob p] [--scheduler greedy|fifo|learned]\n"
                "          [--no-pipeline] [--zerocopy-threshold n] [--span partition|resident]\n"
                "          [--seed n] [--timers] [--json] [--queries file [--tune]] [--pack | --pack-blocks]\n"
                "          [--rowptr-resident] [--split-blocks] [--cache-mb MB]\n",
                app.c_str());
        return 1;
    }
    lt_graph *g = nullptr;
    lt_status s = lt_graph_open(c.file.c_str(), &g);
    if (s != LT_OK) return fail(s, "open graph");
    lt_partition *part = nullptr;
    if (c.materialize) {
        // Config.h:62-66 + EqualSizeStrategy::init's report (PartitionStrategy.h:225-234);
        // --split-blocks: partitions sized for block slots of -s MB instead
        s = c.split_blocks ? lt_partition_build_blocks(g, 1024ull * 1024ull * c.partition_mb, &part)
                           : lt_partition_build(g, 1024ull * 1024ull * c.partition_mb, &part);
        if (s != LT_OK) return fail(s, "partition");
        s = lt_partition_save(part, c.file.c_str());
        if (s != LT_OK) return fail(s, "write .part");
        const uint64_t *rp = lt_graph_rowptr(g);
        uint64_t maxdeg = 0;
        uint32_t maxv = 0;
        for (uint32_t v = 0; v < lt_graph_num_node(g); v++)
            if (rp[v + 1] - rp[v] > maxdeg) {
                maxdeg = rp[v + 1] - rp[v];
                maxv = v;
            }
        printf("node with max degree: %u, degree: %lu\n", maxv, (unsigned long)maxdeg);
    } else {
        s = lt_partition_load(c.file.c_str(), &part);
        if (s != LT_OK) return fail(s, "read .part");
    }
    const uint32_t V = lt_graph_num_node(g);
    const uint32_t P = lt_partition_count(part);
    const uint64_t w = c.has_w ? c.w : 2ull * V;
    // --cache-mb: the raw slots of this table's partition size that fit that many MB
    const uint64_t pcache = c.cache_mb ? std::max<uint64_t>(1, (c.cache_mb << 20) / lt_partition_bytes(part)) : P;
    const uint64_t pdev = c.has_p ? std::min<uint64_t>(c.p, P) : std::min<uint64_t>(pcache, P);
    const uint64_t wdev = c.has_b ? std::min<uint64_t>(c.b, w) : w;

    // batched ppr queries: one "source count" pair per line
    std::vector<uint32_t> qsrc;
    std::vector<uint64_t> qcnt;
    uint64_t total = w;
    if (!c.queries.empty()) {
        FILE *qf = fopen(c.queries.c_str(), "r");
        if (!qf) {
            fprintf(stderr, "cannot open %s\n", c.queries.c_str());
            return 1;
        }
        unsigned long long a, b;
        while (fscanf(qf, "%llu %llu", &a, &b) == 2) {
            qsrc.push_back((uint32_t)a);
            qcnt.push_back(b);
        }
        fclose(qf);
        total = 0;
        for (auto x : qcnt) total += x;
    }
    printf("num nodes: %u, num edges: %lu, num walkers: %lu\n", V, (unsigned long)lt_graph_num_edge(g),
           (unsigned long)total);

    lt_config cfg;
    lt_config_default(&cfg);
    cfg.num_device_partitions = (uint32_t)pdev;
    cfg.num_device_walkers = wdev;
    cfg.seed = c.seed;
    cfg.policy = c.scheduler == "fifo" ? LT_SCHED_FIFO : c.scheduler == "learned" ? LT_SCHED_LEARNED : LT_SCHED_GREEDY;
    cfg.pipeline = c.pipeline ? 1 : 0;
    cfg.span = c.span == "resident" ? LT_SPAN_RESIDENT : LT_SPAN_PARTITION;
    cfg.event_timers = c.timers ? 1 : 0;
    if (c.has_zc) cfg.zerocopy_threshold = c.zc;
    if (c.prefetch_depth >= 0) cfg.prefetch_depth = c.prefetch_depth;
    cfg.zc_batch = c.zc_batch;
    cfg.resident_batch = c.resident_batch;
    cfg.walkers_per_thread = c.walkers_per_thread;
    cfg.slice_chunks = c.slice_chunks;
    cfg.prefetch_plain = c.prefetch_plain;
    cfg.evict_shortlist = c.evict_shortlist;
    cfg.learned_next = c.learned_next;
    cfg.append_mode = c.append_mode;
    cfg.device_schedule = c.device_schedule ? 1 : 0;
    cfg.pack_partitions = c.pack;
    cfg.rowptr_resident = c.rowptr_resident ? 1 : 0;
    if (c.tune && !qsrc.empty()) {
        // the same cache bytes as -p slots of this .part's size, the rest chosen by the autotuner
        lt_tune_trial best;
        const uint64_t cache = std::max<uint64_t>(1, pdev) * lt_partition_bytes(part);
        s = lt_tune_ppr_batch(g, (uint32_t)qsrc.size(), qsrc.data(), qcnt.data(), c.prob, cache, nullptr, 0, 16,
                              &best, nullptr, 0, nullptr);
        if (s != LT_OK) return fail(s, "autotune");
        printf("autotuned: %u MB %s, %u cache slots, span %s, %s scheduling, slice %d, batches of %d, "
               "%s copies, walkers per lane %d\n",
               best.partition_mb, best.split ? "block slots (split for them)" : "partitions", best.slots,
               best.span == LT_SPAN_RESIDENT ? "resident" : "partition",
               best.device_schedule ? "device" : "host", best.slice_chunks, best.zc_batch,
               best.pack_partitions == 2 ? "block" : best.pack_partitions == 1 ? "gap-coded" : "raw",
               best.walkers_per_thread);
        lt_partition_free(part);
        s = best.split ? lt_partition_build_blocks(g, (uint64_t)best.partition_mb << 20, &part)
                       : lt_partition_build(g, (uint64_t)best.partition_mb << 20, &part);
        if (s != LT_OK) return fail(s, "partition");
        cfg.num_device_partitions = best.slots;
        cfg.span = best.span;
        cfg.device_schedule = best.device_schedule;
        cfg.slice_chunks = best.slice_chunks;
        cfg.zc_batch = best.zc_batch;
        cfg.resident_batch = best.zc_batch;
        cfg.pack_partitions = best.pack_partitions;
        cfg.walkers_per_thread = best.walkers_per_thread;
    }
    lt_engine *e = nullptr;
    s = lt_engine_create(g, part, &cfg, &e);
    if (s != LT_OK) return fail(s, "engine");
    lt_scorer *scorer = nullptr;
    if (!c.scorer.empty()) {
        if ((s = lt_scorer_create(part, 1, &scorer)) != LT_OK) return fail(s, "scorer");
        if ((s = lt_scorer_load(scorer, c.scorer.c_str())) != LT_OK) return fail(s, "scorer weights");
        if ((s = lt_engine_set_scorer(e, scorer)) != LT_OK) return fail(s, "scorer");
    }
    printf("number of partitons: %u, graph memory pool: %lu MB, walker memory pool: %lu MB\n",
           lt_partition_count(part),
           (unsigned long)(lt_partition_bytes(part) * cfg.num_device_partitions / 1024 / 1024),
           (unsigned long)(std::min(wdev, total) * 16 / 1024 / 1024));

    std::vector<uint32_t> counts(V, 0);
    if (app == "pagerank") {
        s = lt_run_pagerank(e, w, c.length, (uint32_t)c.runs, counts.data(), nullptr);
    } else if (app == "ppr") {
        if (!qsrc.empty()) {
            std::vector<uint32_t> ends(total);
            s = lt_run_ppr_batch(e, (uint32_t)qsrc.size(), qsrc.data(), qcnt.data(), c.prob, ends.data());
            for (auto v : ends) counts[v]++;
        } else {
            s = lt_run_ppr(e, w, c.source, c.prob, counts.data(), nullptr);
        }
    } else {
        s = lt_run_genericwalk(e, w, c.length, (uint32_t)c.runs, nullptr, nullptr);
    }
    if (s != LT_OK) return fail(s, "run");

    lt_stats st;
    lt_engine_stats(e, &st);
    printf("Running Time: %f ms\n", st.running_ms);
    printf("iterations: %lu, explicit: %lu\n", (unsigned long)st.iterations, (unsigned long)st.explicit_iterations);
    if (!c.timers) printf("Warning: GPU timer is disabled, the following statistics is wrong!\n");
    printf("graph loading time: %f ms, calls: %lu\n", st.graph_loading_ms, (unsigned long)st.graph_loading_calls);
    printf("computing time: %f ms, calls: %lu\n", st.computing_ms, (unsigned long)st.computing_calls);
    printf("zero copy time: %f ms, calls: %lu\n", st.zerocopy_ms, (unsigned long)st.zerocopy_calls);
    printf("walker loading time: %f ms, calls: %lu\n", st.walker_loading_ms, (unsigned long)st.walker_loading_calls);
    printf("walker evicting time: %f ms, calls: %lu\n", st.walker_evicting_ms,
           (unsigned long)st.walker_evicting_calls);
    printf("H2D bytes: explicit %llu, zero-copy %llu\n", (unsigned long long)st.h2d_bytes_explicit,
           (unsigned long long)st.h2d_bytes_zerocopy);
    if (st.h2d_bytes_unpacked != st.h2d_bytes_explicit)
        printf("H2D bytes unpacked: %llu\n", (unsigned long long)st.h2d_bytes_unpacked);
    printf("steps: nominal %llu, moves %llu, teleports %llu\n", (unsigned long long)st.nominal_steps,
           (unsigned long long)st.moves, (unsigned long long)st.teleports);
    printf("throughput: %f steps/s\n", st.running_ms > 0 ? st.nominal_steps / (st.running_ms * 1e-3) : 0.0);

    const uint64_t *rp = lt_graph_rowptr(g);
    int rc = 0;
    if (app == "pagerank") {
        // result(i) and check() of app/pagerank.cu:89-101
        auto result = [&](uint32_t i) {
            return (float)((counts[i] / ((double)w * c.runs / (V + 0.0))) / (c.length + 0.0));
        };
        for (uint32_t i = 0, printed = 0; i < V && printed < 20; i++) {
            const uint64_t deg = rp[i + 1] - rp[i];
            if (deg > 1000) {
                printf("node %u, degree %lu, value %f\n", i, (unsigned long)deg, result(i));
                printed++;
            }
        }
        printf("node 0, degree %lu, value %f\n", (unsigned long)rp[1], result(0));
        size_t sum = 0;
        for (auto x : counts) sum += x;
        const size_t expected = (size_t)w * c.length * c.runs;
        printf("sum: %zu, %zu\n", sum, expected);
        printf(sum == expected ? "result correct\n" : "result incorrect\n");
        if (sum != expected)
            printf("[Warning] ill-conditioned graph such as Yahoo has a very high-degree vertex whose value could "
                   "exceed the max value of uint32.\n");
    } else if (app == "ppr") {
        const uint32_t target = 2;
        if (target < V) printf("PageRank value of node %u: %f\n", target, (float)(counts[target] / (total / (V + 0.0))));
        size_t sum = 0;
        for (auto x : counts) sum += x;
        printf("sum: %lu, %lu\n", (unsigned long)sum, (unsigned long)total);
        printf(sum == total ? "result correct\n" : "result incorrect\n");
    }
    if (c.json) {
        printf("{\"app\": \"%s\", \"running_ms\": %.3f, \"nominal_steps\": %llu, \"moves\": %llu, "
               "\"h2d_bytes_explicit\": %llu, \"h2d_bytes_zerocopy\": %llu, \"iterations\": %llu, "
               "\"explicit\": %llu, \"graph_loads\": %llu}\n",
               app.c_str(), st.running_ms, (unsigned long long)st.nominal_steps, (unsigned long long)st.moves,
               (unsigned long long)st.h2d_bytes_explicit, (unsigned long long)st.h2d_bytes_zerocopy,
               (unsigned long long)st.iterations, (unsigned long long)st.explicit_iterations,
               (unsigned long long)st.graph_loading_calls);
    }
    lt_engine_destroy(e);
    if (scorer) lt_scorer_free(scorer);
    lt_partition_free(part);
    lt_graph_free(g);
    return rc;
}
