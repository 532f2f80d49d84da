// lt_sched.h — partition scheduling policies (host).
//
// GREEDY and FIFO restate Scheduler::nextPartition and the eviction /
// pipelined-compute choices of GPUScheduler (src/Scheduler.h:84-186,
// 368-416).  LEARNED takes GREEDY's choices here; the learned scheduler loop
// (Engine::scheduler_loop_learned) adds its scorer-driven prefetch and victim
// choice on top of them.
#pragma once
#include <stdint.h>

#include <vector>

#include "../../include/lighttraffic.h"

namespace lt {

struct Scheduler {
    uint32_t P = 0;
    int policy = LT_SCHED_GREEDY;
    uint32_t fifo_cursor = 0;  // the reference keeps it in a function-local static (Scheduler.h:394)
    std::vector<int32_t> history;  // loaded partitions, most recent last
    uint64_t learned_decisions = 0, fallback_decisions = 0;  // victims chosen with / without scores

    void init(uint32_t P_, const lt_config &cfg) {
        P = P_;
        policy = cfg.policy;
    }
    void begin_run() {
        fifo_cursor = 0;
        learned_decisions = fallback_decisions = 0;
    }
    void on_load(uint32_t p) { history.push_back((int32_t)p); }

    int32_t next(const uint64_t *counts, int32_t prev, bool sync) {
        if (policy == LT_SCHED_FIFO) return fifo_next(counts);
        return lt_sched_greedy_next(P, counts, prev, sync ? 1 : 0);
    }

    // FIFO nextPartition (src/Scheduler.h:393-416): sweep ids in order,
    // returning -1 at the end of every sweep.
    int32_t fifo_next(const uint64_t *counts) {
        if (fifo_cursor == 0) {
            bool any = false;
            for (uint32_t i = 0; i < P; i++) any |= counts[i] > 0;
            if (!any) return -1;
        }
        while (fifo_cursor < P) {
            const uint32_t c = fifo_cursor++;
            if (counts[c] > 0) return (int32_t)c;
        }
        fifo_cursor = 0;
        return -1;
    }

    int32_t evict_choice(const uint64_t *counts, const uint8_t *resident, int32_t prev_loaded) {
        if (policy == LT_SCHED_FIFO) {
            // round robin after the previously loaded partition (Scheduler.h:178-184)
            const uint32_t start = prev_loaded < 0 ? 0u : (uint32_t)prev_loaded;
            for (uint32_t i = 1; i <= P; i++) {
                const uint32_t v = (start + i) % P;
                if (resident[v] && (int32_t)v != prev_loaded) return (int32_t)v;
            }
            return -1;
        }
        return lt_sched_evict_min(P, counts, resident, prev_loaded);
    }

    int32_t computed_choice(const uint64_t *counts, const uint8_t *resident, int32_t except, uint64_t page) {
        if (policy == LT_SCHED_FIFO) {
            // first resident partition after `except` (Scheduler.h:113-127)
            for (uint32_t i = 1; i < P; i++) {
                const uint32_t v = ((uint32_t)except + i) % P;
                if (resident[v] && counts[v] > 0) return (int32_t)v;
            }
            return -1;
        }
        return lt_sched_computed_partition(P, counts, resident, except, page);
    }
};

}  // namespace lt
