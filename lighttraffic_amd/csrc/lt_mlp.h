// lt_mlp.h — the learned partition scorer (host interface, no HIP types).
//
// A per-partition MLP (16 features -> 256 -> 256 -> 2, BF16 weights, FP32
// accumulation) evaluated on all P partitions of one scheduler decision.
// Output 0 ("next") scores how likely a partition is to be the one the
// scheduler picks at its next decision; output 1 ("keep") predicts how soon
// the partition is next walked from HBM (or, with the older labels, its share
// of future walker traffic).  The engine prefetches the best "next" partition
// while the current one is walked and, among the resident partitions with the
// fewest queued walkers, evicts the one with the lowest "keep".
#pragma once
#include <stdint.h>

#include <vector>

#include "../../include/lighttraffic.h"

namespace lt {

constexpr int kFeat = 16;
constexpr int kHidden = 256;
constexpr int kOut = 2;
constexpr int kRecent = 8;

// One scheduler decision as the scorer sees it.
struct DecisionState {
    const uint32_t *counts;    // [P] walkers queued per partition
    const uint8_t *resident;   // [P] 1 if the partition is in the HBM cache
    const uint32_t *age;       // [P] HBM walks of any partition since this one's last
    int32_t recent[kRecent];   // last loaded partitions, most recent first, -1 = none
    int32_t cur;               // partition chosen at this decision, -1 = none
    float w_total;             // walkers of the run
};

// Training samples accumulated from traces (host memory), one per decision.
struct Samples {
    uint64_t n = 0;
    std::vector<uint32_t> counts;   // [n*P]
    std::vector<uint8_t> resident;  // [n*P]
    std::vector<uint32_t> age;      // [n*P]
    std::vector<int32_t> recent;    // [n*kRecent]
    std::vector<int32_t> cur, next; // [n]
    std::vector<float> w_total, wa, wb;  // [n]
    std::vector<float> keep;        // [n*P]
};

struct Scorer;  // defined in lt_mlp.hip

Samples &scorer_samples(Scorer *s);
uint32_t scorer_partitions(const Scorer *s);
Scorer *scorer_new(uint32_t P, const uint32_t *node_off, const uint64_t *edge_off, uint64_t seed, lt_status *st);
void scorer_free(Scorer *s);
bool scorer_trained(const Scorer *s);
// Scores of one decision: out[P*2] on the host.  Runs on `stream` (a
// hipStream_t passed as void*, nullptr = the scorer's own stream) and
// synchronises with it.
lt_status scorer_infer(Scorer *s, const DecisionState &d, void *stream, float *out);
// The same in two halves: launch the inference of one decision (the state is
// copied at once), then wait for its scores.  One inference in flight per
// scorer; a launch waits for the previous one's kernel first.
lt_status scorer_infer_launch(Scorer *s, const DecisionState &d, void *stream);
lt_status scorer_infer_wait(Scorer *s, float *out);
bool scorer_infer_pending(const Scorer *s);

}  // namespace lt
