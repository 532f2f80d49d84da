// lt_engine.hip — the GPU engine behind lt_engine_* / lt_run_*.
//
// Host side of the hot path: the HBM partition cache (role of GPUGraph,
// src/GraphLoader.h:160-240), the walker queues (GPUListPool,
// src/buffer/Pool.h:530-646), the per-partition processing step
// (GPUWalkManager::processInMemory, src/WalkManager.h:362-404) and the plain
// scheduler loop (GPUScheduler / Scheduler, src/Scheduler.h:13-454).
#include <hip/hip_runtime.h>

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <memory>
#include <vector>

#include "lt_internal.h"
#include "lt_kernels.h"
#include "lt_mlp.h"
#include "lt_sched.h"

#define LT_HIP(call)                                                                                   \
    do {                                                                                               \
        hipError_t _e = (call);                                                                        \
        if (_e != hipSuccess) {                                                                        \
            lt::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(_e));            \
            return LT_E_HIP;                                                                           \
        }                                                                                              \
    } while (0)

#define LT_TRY(expr)                    \
    do {                                \
        lt_status _s = (expr);          \
        if (_s != LT_OK) return _s;     \
    } while (0)

namespace lt {

// Reference page size (numBlock * threadPerBlock * walkerPerThread,
// src/Type.h:10-14): the "full queue" threshold of getComputedPartition.
constexpr uint64_t kRefPage = 82ull * 512 * 4;

template <typename T>
static lt_status dmalloc(T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, sizeof(T) * count);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        set_error("hipMalloc of %zu bytes failed: %s", sizeof(T) * count, hipGetErrorString(e));
        return LT_E_NOMEM;
    }
    return LT_OK;
}

template <typename T>
static void dfree(T *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

// Accumulates HIP-event intervals on one stream (GPUTimer, src/utils/Stopwatch.h:49-104).
struct EventTimer {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> open;
    std::vector<hipEvent_t> pool;
    double ms = 0;
    uint64_t calls = 0;
    bool enabled = false;
    hipEvent_t take() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        (void)hipEventCreate(&e);
        return e;
    }
    hipEvent_t begin(hipStream_t s) {
        calls++;
        if (!enabled) return nullptr;
        hipEvent_t a = take();
        (void)hipEventRecord(a, s);
        return a;
    }
    void end(hipEvent_t a, hipStream_t s) {
        if (!enabled || !a) return;
        hipEvent_t b = take();
        (void)hipEventRecord(b, s);
        open.emplace_back(a, b);
    }
    void collect() {
        for (auto &pr : open) {
            float t = 0;
            if (hipEventSynchronize(pr.second) == hipSuccess && hipEventElapsedTime(&t, pr.first, pr.second) == hipSuccess)
                ms += t;
            pool.push_back(pr.first);
            pool.push_back(pr.second);
        }
        open.clear();
    }
    void reset() {
        collect();
        ms = 0;
        calls = 0;
    }
    ~EventTimer() {
        collect();
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

// Waits on the scheduler's critical path poll instead of blocking: every
// decision waits for the fix-up of the previous walk, and a blocking wait adds
// its wake-up latency to each of them.
static hipError_t spin_stream(hipStream_t st) {
    hipError_t e;
    while ((e = hipStreamQuery(st)) == hipErrorNotReady) {
    }
    return e;
}

static hipError_t spin_event(hipEvent_t ev) {
    hipError_t e;
    while ((e = hipEventQuery(ev)) == hipErrorNotReady) {
    }
    return e;
}

struct RunSpec {
    int app = APP_PAGERANK;
    uint64_t total = 0;     // walker ids [0, total)
    uint32_t length = 0;    // pagerank / genericwalk l
    uint32_t source = 0;    // ppr
    float prob = 0.15f;     // ppr
    uint64_t nominal = 0;   // nominal steps for the stats
    // batched ppr
    const uint32_t *q_src = nullptr;
    const uint64_t *q_cnt = nullptr;
    uint32_t nq = 0;
    uint32_t *paths_out = nullptr;  // genericwalk: u32[total][length + 1] on the host
};

struct Engine {
    const lt_graph *g = nullptr;
    lt_partition part;
    lt_config cfg;
    int device = 0;
    uint32_t V = 0, P = 0;
    hipStream_t s_compute = nullptr, s_copy = nullptr;
    int num_cu = 0;
    int max_grid = 0;           // init kernel (4 walkers per thread)
    int walk_grid[3][5] = {};  // resident workgroups of each walk kernel, by app and walkers per thread

    // partition table
    uint32_t *d_node_off = nullptr;
    uint64_t *d_edge_off = nullptr;
    uint32_t *d_lut = nullptr;
    uint32_t lut_shift = 0;
    const uint64_t **d_rowptr_tab = nullptr;
    const uint32_t **d_col_tab = nullptr;
    const uint32_t **d_words_tab = nullptr;  // packed slots (pack_partitions = 2): bit streams

    // HBM partition cache
    uint32_t nslots = 0;
    uint64_t slot_sz = 0;        // bytes per slot (256-aligned)
    bool slot_packed = false;    // slots keep columns packed (pack_partitions = 2)
    uint8_t *d_slots = nullptr;
    std::vector<int32_t> slot_of;      // partition -> slot or -1
    std::vector<int32_t> part_in_slot; // slot -> partition or -1
    std::vector<uint32_t> free_slots;
    std::vector<uint8_t> resident;     // partition -> 0/1
    uint64_t zc_threshold = 0;
    const uint64_t *zc_rowptr = nullptr;  // device view of the pinned host graph
    const uint32_t *zc_col = nullptr;
    // Packed copies: images land in two HBM staging buffers by DMA on s_dma;
    // s_copy waits for the DMA, then unpacks into the slot, so the next DMA
    // overlaps this unpack.  Loads are complete when s_copy reaches them.
    // With too few hardware queues s_dma is s_copy (DMA and unpack in turn).
    std::shared_ptr<const PackedImage> pk_hold;  // packed partition images (pack_partitions)
    const PackedImage *pk = nullptr;
    uint8_t *d_stage[2] = {nullptr, nullptr};
    hipEvent_t stage_free[2] = {nullptr, nullptr};  // unpack of the buffer's last image done
    bool stage_used[2] = {false, false};
    uint32_t stage_next = 0;
    hipStream_t s_dma = nullptr;          // == s_copy without packed copies
    hipEvent_t ev_dma = nullptr;          // last DMA queued on s_dma

    // walker queues
    uint64_t wcap = 0;  // walkers the pool is sized for
    Queues q{};
    uint32_t *d_meta = nullptr;   // [P] qlen followed by PoolCtl
    uint32_t *h_meta = nullptr;   // pinned mirror
    std::vector<uint64_t> counts; // walkers queued per partition: device queue + spilled to host
    std::vector<uint64_t> dcounts;  // device part of counts (qlen)

    // Host walker spill (src/WalkManager.h:11-58, src/Scheduler.h:53-82,
    // 215-233): with more walkers than num_device_walkers, whole device
    // queues move to per-partition queues in pinned, mapped host memory, and
    // come back when their partition is walked (the kernel reads them over
    // PCIe).  Host chunks are kChunk walkers, handed out from pinned blocks.
    static constexpr uint32_t kHostBlockChunks = 256;  // 16 MiB per pinned block
    bool spill_on = false;
    uint64_t dev_cap = 0;                   // walkers the device queues may hold
    std::vector<Walker *> hblocks, hblocks_dev;
    std::vector<uint32_t> hfree;            // free host chunk ids
    std::vector<std::deque<uint32_t>> hq;   // per partition: host chunks, oldest first
    std::vector<uint64_t> hhead, hcount;    // first walker's slot in hq[p].front(); walkers
    uint64_t spill_h2d_bytes = 0;           // spilled walkers read back over PCIe
    Walker *host_chunk_ptr(uint32_t id, bool dev) const {
        return (dev ? hblocks_dev : hblocks)[id / kHostBlockChunks] + (size_t)(id % kHostBlockChunks) * kChunk;
    }
    lt_status host_chunk(uint32_t &id);
    void host_reset();
    int32_t spill_victim(uint32_t except) const;
    lt_status spill_round(const std::vector<uint32_t> &victims);
    uint8_t *d_plan = nullptr, *h_plan = nullptr;  // spill round plan (device copy, pinned staging)
    size_t plan_cap = 0;
    lt_status ensure_room(uint64_t want, uint32_t except, uint64_t &room);
    lt_status process_host(uint32_t p, const WalkArgs &a);

    // outputs and statistics
    uint32_t *d_counts = nullptr;
    size_t counts_cap = 0;
    uint32_t count_copies = 1;  // 1 or kCountCopies arrays of V counts
    uint32_t *d_endpoints = nullptr;
    size_t endpoints_cap = 0;
    uint32_t *d_moves = nullptr;
    uint32_t *d_paths = nullptr;
    size_t paths_cap = 0;
    size_t moves_cap = 0;
    uint64_t *d_qoff = nullptr;
    uint32_t *d_qsrc = nullptr;
    size_t q_cap = 0;
    unsigned long long *d_stats = nullptr;

    lt_stats st{};
    std::vector<lt_trace_rec> trace;
    std::vector<uint32_t> trace_counts;  // queue lengths seen by each traced decision
    std::vector<uint8_t> trace_resident; // residency seen by each traced decision
    std::vector<int32_t> trace_recent;   // last kRecent loads at each traced decision
    std::vector<uint32_t> trace_age;     // partition ages seen by each traced decision
    // Recency of HBM use, a scorer feature: use_clock counts the walks of
    // resident partitions since the engine was created, last_use[p] is the
    // clock after p's last one (0 = never), and p's age is their difference.
    std::vector<uint64_t> last_use;
    uint64_t use_clock = 0;
    Scorer *scorer = nullptr;             // learned policy (not owned)
    // device-side scheduling state (mirrors of slot_of / part_in_slot / resident)
    DevSched *d_ds = nullptr;
    int32_t *d_slot_of = nullptr, *d_part_in_slot = nullptr;
    uint8_t *d_resident = nullptr;
    uint8_t *d_zc_part = nullptr;  // [P] partitions of the running zero-copy batch
    bool rp_resident = false;      // rowptr of the whole graph kept in HBM (zc_rowptr points there)
    uint64_t *d_pk_off = nullptr, *d_pk_size = nullptr;  // device loop with block slots: image table
    const uint8_t *d_pk_img = nullptr;                     // device view of the pinned images
    uint64_t *d_rowptr_all = nullptr;
    uint64_t pending_h2d = 0;      // one-off H2D bytes not yet charged to a run
    uint32_t zc_batch = 1;         // partitions per zero-copy launch
    uint32_t res_batch = 1;        // partitions per launch over resident queues
    uint32_t slice = kSliceChunks;
    std::vector<hipEvent_t> pool_events;  // reusable prefetch-completion events
    std::vector<float> scores;            // scorer output of the current decision
    bool scores_valid = false;
    uint64_t scores_clock = 0;            // use_clock when the scores were computed
    bool infer_pending = false;           // an inference launched by infer_launch, not yet collected
    uint64_t pending_clock = 0;           // use_clock at its launch
    uint64_t w_run = 0;                   // walkers of the current wave
    EventTimer t_load, t_compute, t_zc, t_wload, t_wevict, t_kernel, t_kzc;  // t_kzc: walk launches over PCIe
    Scheduler sched;

    ~Engine();
    lt_status init(const lt_graph *g_, const lt_partition *p_, const lt_config *c);
    lt_status ensure_pool(uint64_t walkers);
    lt_status ensure_outputs(const RunSpec &rs, bool want_endpoints, bool want_moves);
    lt_status run(const RunSpec &rs, uint32_t *counts_out, uint32_t *endpoints_out, uint32_t *moves_out);

    // scheduler primitives
    lt_status read_meta();
    lt_status take_event(hipEvent_t &ev) {
        if (!pool_events.empty()) {
            ev = pool_events.back();
            pool_events.pop_back();
            return LT_OK;
        }
        LT_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        return LT_OK;
    }
    bool meta_dirty = true;  // q.packed may be older than the queues
    hipEvent_t ev_evicted = nullptr;  // table updates of prefetch evictions, awaited by the copy stream
    lt_status process(uint32_t p, bool zerocopy, uint64_t walkers_hint);
    lt_status process_zc(uint32_t c, const uint8_t *pending);
    lt_status walk_batch(const std::vector<uint32_t> &parts, bool zerocopy);
    lt_status process_resident(uint32_t c, int kind, uint8_t *prefetched);
    void record(uint32_t p, int kind, int32_t loaded, int32_t evicted);
    lt_status evict(uint32_t p);
    lt_status load_async(uint32_t p);
    lt_status copy_partition(uint32_t p, uint8_t *base);
    lt_status fetch_partition(uint32_t p, uint64_t *rowptr_out, uint32_t *col_out);
    // Bit streams of p's slot whose column part starts at cp: past the block
    // headers when the slot holds p's block image, nullptr for raw columns.
    const uint32_t *slot_words(uint32_t p, const uint32_t *cp) const {
        if (!slot_packed || !pk->size[p]) return nullptr;
        const uint64_t nblk = (part.edge_off[p + 1] - part.edge_off[p] + kBlockEdges - 1) / kBlockEdges;
        return cp + 2 * (nblk + 1);
    }
    lt_status set_table(uint32_t p, bool present);
    lt_status scheduler_loop();
    lt_status scheduler_loop_learned();
    lt_status scheduler_loop_device();
    SchedState sched_state(bool device) const;
    lt_status infer_scores(int32_t cur);
    lt_status infer_launch(int32_t cur);
    lt_status infer_collect();
    int32_t learned_victim(int32_t a, int32_t b);
    void recent_loads(int32_t *out) const;
    uint32_t age_of(uint32_t p) const;
    // scores at most kStaleWalks walks of resident partitions old (the ages
    // and residency they saw have moved by at most that much)
    static constexpr uint64_t kStaleWalks = 16;
    bool scores_recent() const { return scores_valid && use_clock - scores_clock <= kStaleWalks; }
    lt_status run_partition(uint32_t p, int32_t &prev_loaded);

    AppArgs app{};
    int app_kind = APP_PAGERANK;
    size_t lds_bytes(int wpt = kWpt) const { return append_lds_bytes(P, wpt); }
    lt_status launch_walk(const WalkArgs &a, uint64_t walkers, hipStream_t s);
    uint64_t partition_bytes(uint32_t p) const {
        return ((uint64_t)part.node_off[p + 1] - part.node_off[p] + 1) * 8 +
               (part.edge_off[p + 1] - part.edge_off[p]) * 4;
    }
};

Engine::~Engine() {
    if (s_compute) (void)hipStreamSynchronize(s_compute);
    if (s_copy) (void)hipStreamSynchronize(s_copy);
    if (s_dma && s_dma != s_copy) (void)hipStreamSynchronize(s_dma);
    for (auto ev : pool_events) (void)hipEventDestroy(ev);
    if (ev_evicted) (void)hipEventDestroy(ev_evicted);
    dfree(d_node_off);
    dfree(d_edge_off);
    dfree(d_lut);
    dfree(d_rowptr_tab);
    dfree(d_col_tab);
    dfree(d_words_tab);
    dfree(d_slots);
    dfree(q.slab);
    dfree(q.frame);
    dfree(q.free_stack);
    dfree(q.lost);
    dfree(d_meta);
    if (h_meta) (void)hipHostFree(h_meta);
    for (Walker *b : hblocks) (void)hipHostFree(b);
    dfree(d_plan);
    if (h_plan) (void)hipHostFree(h_plan);
    dfree(d_counts);
    dfree(d_endpoints);
    dfree(d_moves);
    dfree(d_paths);
    dfree(d_qoff);
    dfree(d_qsrc);
    dfree(d_stats);
    dfree(d_ds);
    dfree(d_slot_of);
    dfree(d_part_in_slot);
    dfree(d_resident);
    dfree(d_zc_part);
    dfree(d_rowptr_all);
    dfree(d_pk_off);
    dfree(d_pk_size);
    dfree(d_stage[0]);
    dfree(d_stage[1]);
    if (s_compute) (void)hipStreamDestroy(s_compute);
    if (s_copy) (void)hipStreamDestroy(s_copy);
    if (s_dma && s_dma != s_copy) (void)hipStreamDestroy(s_dma);
    for (int k = 0; k < 2; k++)
        if (stage_free[k]) (void)hipEventDestroy(stage_free[k]);
    if (ev_dma) (void)hipEventDestroy(ev_dma);
}

lt_status Engine::init(const lt_graph *g_, const lt_partition *p_, const lt_config *c) {
    g = g_;
    part = *p_;
    cfg = *c;
    V = g->V;
    P = part.P();
    LT_TRY(check_partition_of(*g, part));
    if (P > kMaxPartitions) {
        set_error("%u partitions exceed the engine limit %u; use a larger partition size", P, kMaxPartitions);
        return LT_E_ARG;
    }
    if (V == 0) {
        set_error("empty graph");
        return LT_E_ARG;
    }
    if (cfg.evict_shortlist < 0) {
        set_error("evict_shortlist must be >= 0 (got %d)", cfg.evict_shortlist);
        return LT_E_ARG;
    }
    if (cfg.append_mode < 0 || cfg.append_mode > 2) {
        set_error("append_mode must be 0, 1 or 2 (got %d)", cfg.append_mode);
        return LT_E_ARG;
    }
    if (cfg.learned_next < 0 || cfg.learned_next > 100) {
        set_error("learned_next must be 0-100 (got %d)", cfg.learned_next);
        return LT_E_ARG;
    }
    if (cfg.prefetch_plain < 0 || cfg.prefetch_plain > 2) {
        set_error("prefetch_plain must be 0, 1 or 2 (got %d)", cfg.prefetch_plain);
        return LT_E_ARG;
    }
    if (cfg.walkers_per_thread != 0 && cfg.walkers_per_thread != 1 && cfg.walkers_per_thread != 2 &&
        cfg.walkers_per_thread != 4) {
        set_error("walkers_per_thread must be 0, 1, 2 or 4 (got %d)", cfg.walkers_per_thread);
        return LT_E_ARG;
    }
    if (cfg.pack_partitions < 0 || cfg.pack_partitions > 2) {
        set_error("pack_partitions must be 0, 1 or 2 (got %d)", cfg.pack_partitions);
        return LT_E_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        set_error("no HIP device visible");
        return LT_E_NODEVICE;
    }
    if (!g->pinned) {
        set_error("graph was loaded without a HIP device; reopen it");
        return LT_E_ARG;
    }
    LT_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    LT_HIP(hipGetDeviceProperties(&prop, device));
    num_cu = prop.multiProcessorCount;
    LT_HIP(hipStreamCreateWithFlags(&s_compute, hipStreamNonBlocking));
    LT_HIP(hipStreamCreateWithFlags(&s_copy, hipStreamNonBlocking));
    int occ = 0;
    LT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, init_kernel<APP_PAGERANK>, kBlock, lds_bytes()));
    max_grid = std::max(1, occ) * num_cu;
    // Each app's kernels have their own register counts, so their own
    // occupancy: a grid sized by another kernel's would queue workgroups
    // behind the resident ones and finish in a second, partial round.
    const void *wk[3][5] = {
        {nullptr, (const void *)walk_kernel<APP_PAGERANK, 1>, (const void *)walk_kernel<APP_PAGERANK, 2>, nullptr,
         (const void *)walk_kernel<APP_PAGERANK, 4>},
        {nullptr, (const void *)walk_kernel<APP_PPR, 1>, (const void *)walk_kernel<APP_PPR, 2>, nullptr,
         (const void *)walk_kernel<APP_PPR, 4>},
        {nullptr, (const void *)walk_kernel<APP_GENERIC, 1>, (const void *)walk_kernel<APP_GENERIC, 2>, nullptr,
         (const void *)walk_kernel<APP_GENERIC, 4>}};
    for (int app : {APP_PAGERANK, APP_PPR, APP_GENERIC})
        for (int wpt : {1, 2, 4}) {
            LT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, wk[app][wpt], kBlock, lds_bytes(wpt)));
            walk_grid[app][wpt] = std::max(1, occ) * num_cu;
        }

    // partition table + vertex -> partition lookup table
    LT_TRY(dmalloc(&d_node_off, P + 1));
    LT_TRY(dmalloc(&d_edge_off, P + 1));
    LT_HIP(hipMemcpy(d_node_off, part.node_off.data(), 4 * (P + 1), hipMemcpyHostToDevice));
    LT_HIP(hipMemcpy(d_edge_off, part.edge_off.data(), 8 * (P + 1), hipMemcpyHostToDevice));
    uint32_t bits = 0;
    while (bits < 32 && (1ull << bits) < (uint64_t)V) bits++;
    lut_shift = bits > 16 ? bits - 16 : 0;
    const size_t nlut = (((size_t)V - 1) >> lut_shift) + 1;
    std::vector<uint32_t> lut(nlut);
    for (size_t j = 0; j < nlut; j++) lut[j] = part.find((uint32_t)(j << lut_shift));
    LT_TRY(dmalloc(&d_lut, nlut));
    LT_HIP(hipMemcpy(d_lut, lut.data(), 4 * nlut, hipMemcpyHostToDevice));
    LT_TRY(dmalloc(&d_rowptr_tab, P));
    LT_TRY(dmalloc(&d_col_tab, P));
    LT_HIP(hipMemset(d_rowptr_tab, 0, sizeof(void *) * P));
    LT_HIP(hipMemset(d_col_tab, 0, sizeof(void *) * P));
    LT_TRY(dmalloc(&d_words_tab, P));
    LT_HIP(hipMemset(d_words_tab, 0, sizeof(void *) * P));

    // Packed copies: the images are built once per graph and partition table.
    // Format 1 travels gap-coded and is unpacked into raw slots; format 2 is
    // read in place, so slots are as large as the largest packed partition and
    // the same cache bytes hold more of them.
    s_dma = s_copy;
    // The device loop copies with load_kernel and has no unpack step: it
    // takes block slots (2), not gap-coded copies (1).
    if (cfg.pack_partitions == 2 || (cfg.pack_partitions == 1 && !cfg.device_schedule)) {
        LT_TRY(packed_image(*g, part, cfg.pack_partitions >= 2 ? 2 : 1, &pk_hold));
        pk = pk_hold.get();
        if (!pk->max_size) {
            pk = nullptr;  // no partition gets smaller
            pk_hold.reset();
        }
    }
    slot_packed = pk && pk->format == 2;

    // HBM cache slots (GPUGraph, src/GraphLoader.h:175-176)
    uint64_t slot_bytes = 0;
    slot_layout(part, slot_packed ? pk->size.data() : nullptr, cfg.num_device_partitions, &slot_bytes, &nslots);
    slot_sz = slot_bytes;
    // The whole rowptr array in HBM, paid for with slots of the same cache:
    // a zero-copied move then reads only its column entry over PCIe.
    const uint64_t rp_bytes = 8ull * ((uint64_t)V + 1);
    const uint32_t rp_slots = (uint32_t)((rp_bytes + slot_bytes - 1) / slot_bytes);
    // Off unless asked for: on R-MAT s26 pagerank with an 8 GB cache the
    // slots it takes cost about the PCIe bytes it saves at the default
    // threshold.  At least two slots stay, so that eviction, which spares the
    // partition loaded last, has a victim.
    rp_resident = cfg.rowptr_resident > 0;
    if (nslots == P) rp_resident = false;  // nothing is zero-copied with every partition cached
    if (rp_resident && nslots >= rp_slots + 2) nslots -= rp_slots;
    else rp_resident = false;
    LT_TRY(dmalloc(&d_slots, (size_t)nslots * slot_bytes));
    slot_of.assign(P, -1);
    part_in_slot.assign(nslots, -1);
    resident.assign(P, 0);
    last_use.assign(P, 0);
    free_slots.clear();
    for (uint32_t s = 0; s < nslots; s++) free_slots.push_back(nslots - 1 - s);
    // Default threshold: the reference's partitionSize / cachelineSize / 2
    // (src/Scheduler.h:295) for the plain policies.  The learned policy uses
    // the byte break-even instead: a zero-copied walker step reads a rowptr
    // pair and a column id, two 64-byte PCIe reads, so reading S bytes by DMA
    // costs the same as S/128 zero-copied walkers.
    if (cfg.zerocopy_threshold != UINT64_MAX) zc_threshold = cfg.zerocopy_threshold;
    else zc_threshold = cfg.policy == LT_SCHED_LEARNED ? part.bytes / 128 : part.bytes / 128 / 2;
    LT_HIP(hipHostGetDevicePointer((void **)&zc_rowptr, g->rowptr, 0));
    LT_HIP(hipHostGetDevicePointer((void **)&zc_col, g->col, 0));
    if (pk && pk->format == 1) {  // staging buffers for the images unpacked in HBM
        for (int k = 0; k < 2; k++) {
            LT_TRY(dmalloc(&d_stage[k], pk->max_size));
            LT_HIP(hipEventCreateWithFlags(&stage_free[k], hipEventDisableTiming));
        }
        // HIP spreads streams over GPU_MAX_HW_QUEUES hardware queues (4
        // unless set).  A DMA stream that shares a queue with the walk
        // stream waits behind walk kernels, so it is used only when the
        // streams in play (null, walk, copy, DMA, a scorer's) get their own.
        const char *hwq = getenv("GPU_MAX_HW_QUEUES");
        if (hwq && atoi(hwq) >= 6) {
            LT_HIP(hipEventCreateWithFlags(&ev_dma, hipEventDisableTiming));
            LT_HIP(hipStreamCreateWithFlags(&s_dma, hipStreamNonBlocking));
        }
    }
    if (cfg.device_schedule && slot_packed) {  // the device loop's copies read the images in place
        std::vector<uint64_t> off(P), size(P);
        for (uint32_t i = 0; i < P; i++) {
            off[i] = pk->off[i];
            size[i] = pk->size[i];
        }
        LT_TRY(dmalloc(&d_pk_off, P));
        LT_TRY(dmalloc(&d_pk_size, P));
        LT_HIP(hipMemcpy(d_pk_off, off.data(), 8 * (size_t)P, hipMemcpyHostToDevice));
        LT_HIP(hipMemcpy(d_pk_size, size.data(), 8 * (size_t)P, hipMemcpyHostToDevice));
        LT_HIP(hipHostGetDevicePointer((void **)&d_pk_img, pk->host, 0));
    }
    if (rp_resident) {
        LT_TRY(dmalloc(&d_rowptr_all, (size_t)V + 1));
        LT_HIP(hipMemcpy(d_rowptr_all, g->rowptr, rp_bytes, hipMemcpyHostToDevice));
        pending_h2d = rp_bytes;  // charged to the first run
        zc_rowptr = d_rowptr_all;
    }

    // queue lengths one 128-byte line apart, then PoolCtl on its own line; the
    // fix-up kernel packs both into h_meta (pinned, mapped) for the host
    LT_TRY(dmalloc(&d_meta, (size_t)P * kQStride + kQStride));
    LT_HIP(hipHostMalloc((void **)&h_meta, 4 * (P + 4), hipHostMallocMapped | hipHostMallocCoherent));
    LT_HIP(hipHostGetDevicePointer((void **)&q.packed, h_meta, 0));
    q.qlen = d_meta;
    q.ctl = (PoolCtl *)(d_meta + (size_t)P * kQStride);
    q.P = P;
    LT_TRY(dmalloc(&d_ds, 1));
    LT_TRY(dmalloc(&d_slot_of, P));
    LT_TRY(dmalloc(&d_part_in_slot, nslots));
    LT_TRY(dmalloc(&d_resident, P));
    LT_TRY(dmalloc(&d_zc_part, P));
    LT_HIP(hipMemset(d_zc_part, 0, P));
    zc_batch = cfg.zc_batch > 0 ? (uint32_t)cfg.zc_batch
                                : (cfg.policy == LT_SCHED_LEARNED && cfg.span == LT_SPAN_RESIDENT ? kMaxZcBatch : 1);
    zc_batch = std::min(zc_batch, kMaxZcBatch);
    if (cfg.span != LT_SPAN_RESIDENT) zc_batch = 1;  // a batch relies on walking through the table
    res_batch = cfg.resident_batch > 0
                    ? (uint32_t)cfg.resident_batch
                    : (cfg.policy == LT_SCHED_LEARNED && cfg.span == LT_SPAN_RESIDENT ? kMaxZcBatch : 1);
    res_batch = std::min(res_batch, kMaxZcBatch);
    if (cfg.span != LT_SPAN_RESIDENT) res_batch = 1;
    slice = cfg.slice_chunks > 0 ? (uint32_t)cfg.slice_chunks : kSliceChunks;
    LT_TRY(dmalloc(&d_stats, (size_t)kNumStats * kStatSlots));
    counts.assign(P, 0);
    dcounts.assign(P, 0);
    hq.assign(P, {});
    hhead.assign(P, 0);
    hcount.assign(P, 0);

    bool timers = cfg.event_timers != 0;
    t_load.enabled = t_compute.enabled = t_zc.enabled = t_wload.enabled = t_wevict.enabled = timers;
    t_kernel.enabled = t_kzc.enabled = timers;
    sched.init(P, cfg);
    return LT_OK;
}

// Size the chunk pool for `walkers` queued walkers: every queue may hold a
// partly filled chunk, the slice being consumed keeps its chunks until the
// fix-up, and allocation races may strand up to kLostCap chunks per launch.
lt_status Engine::ensure_pool(uint64_t walkers) {
    if (walkers <= wcap && q.slab) return LT_OK;
    dfree(q.slab);
    dfree(q.frame);
    dfree(q.free_stack);
    dfree(q.lost);
    const uint64_t wchunks = (walkers + kChunk - 1) / kChunk;
    // one launch holds at most `held` source chunks while filling new ones
    const uint64_t held = cfg.device_schedule ? wchunks + 1 : std::min<uint64_t>(slice, wchunks + 1);
    const uint64_t nchunks = wchunks + P + held + 2 + kLostCap;
    const uint64_t maxc = wchunks + 2;
    if (nchunks >= 0xFFFFFFF0ull || maxc * kChunk > 0xFFFFFFFFull) {
        set_error("%llu walkers exceed the queue limits; lower -b", (unsigned long long)walkers);
        return LT_E_ARG;
    }
    q.nchunks = (uint32_t)nchunks;
    q.maxc = (uint32_t)maxc;
    LT_TRY(dmalloc(&q.slab, (size_t)nchunks * kChunk));
    LT_TRY(dmalloc(&q.frame, (size_t)P * maxc));
    LT_TRY(dmalloc(&q.free_stack, nchunks));
    LT_TRY(dmalloc(&q.lost, kLostCap));
    wcap = walkers;
    return LT_OK;
}

lt_status Engine::ensure_outputs(const RunSpec &rs, bool want_endpoints, bool want_moves) {
    const size_t ncount = (size_t)V * count_copies;
    if (counts_cap < ncount) {
        dfree(d_counts);
        LT_TRY(dmalloc(&d_counts, ncount));
        counts_cap = ncount;
    }
    if (want_endpoints && endpoints_cap < rs.total) {
        dfree(d_endpoints);
        LT_TRY(dmalloc(&d_endpoints, rs.total));
        endpoints_cap = rs.total;
    }
    if (rs.paths_out) {
        const size_t n = (size_t)rs.total * ((size_t)rs.length + 1);
        if (paths_cap < n) {
            dfree(d_paths);
            LT_TRY(dmalloc(&d_paths, n));
            paths_cap = n;
        }
    }
    if (want_moves && moves_cap < rs.total) {
        dfree(d_moves);
        LT_TRY(dmalloc(&d_moves, rs.total));
        moves_cap = rs.total;
    }
    if (rs.nq && q_cap < rs.nq) {
        dfree(d_qoff);
        dfree(d_qsrc);
        LT_TRY(dmalloc(&d_qoff, rs.nq + 1));
        LT_TRY(dmalloc(&d_qsrc, rs.nq));
        q_cap = rs.nq;
    }
    return LT_OK;
}

lt_status Engine::read_meta() {
    if (meta_dirty) {  // the last queue update on s_compute was not followed by a fix-up
        pack_meta_kernel<<<1, kBlock, 0, s_compute>>>(q);
        LT_HIP(hipGetLastError());
        meta_dirty = false;
    }
    LT_HIP(spin_stream(s_compute));  // the last fix-up wrote h_meta
    const PoolCtl *c = (const PoolCtl *)(h_meta + P);
    if (c->overflow) {
        set_error("walker queue pool overflow (flags 0x%x, top %u)", c->overflow, c->top);
        return LT_E_INTERNAL;
    }
    for (uint32_t i = 0; i < P; i++) {
        dcounts[i] = h_meta[i];
        counts[i] = dcounts[i] + hcount[i];
    }
    return LT_OK;
}

lt_status Engine::set_table(uint32_t p, bool present) {
    const uint64_t *rp = nullptr;
    const uint32_t *cp = nullptr, *wp = nullptr;
    if (present) {
        uint8_t *base = d_slots + (size_t)slot_of[p] * slot_sz;
        rp = (const uint64_t *)base;
        cp = (const uint32_t *)(base + ((uint64_t)part.node_off[p + 1] - part.node_off[p] + 1) * 8);
        wp = slot_words(p, cp);
    }
    set_table_kernel<<<1, 64, 0, s_compute>>>(d_rowptr_tab, d_col_tab, d_words_tab, p, rp, cp, wp);
    LT_HIP(hipGetLastError());
    return LT_OK;
}

lt_status Engine::evict(uint32_t p) {
    const int32_t s = slot_of[p];
    if (s < 0) return LT_OK;
    LT_TRY(set_table(p, false));
    slot_of[p] = -1;
    part_in_slot[s] = -1;
    resident[p] = 0;
    free_slots.push_back((uint32_t)s);
    return LT_OK;
}

// GPUGraph::loadAsync (src/GraphLoader.h:193-215): rowptr slice at slot offset
// 0, col slice right after it, both on the copy stream.  The compute stream
// must not be reading the slot (the caller evicted it after a sync).
lt_status Engine::load_async(uint32_t p) {
    if (free_slots.empty()) {
        set_error("no free cache slot for partition %u", p);
        return LT_E_INTERNAL;
    }
    const uint32_t s = free_slots.back();
    free_slots.pop_back();
    hipEvent_t t = t_load.begin(s_dma);
    LT_TRY(copy_partition(p, d_slots + (size_t)s * slot_sz));
    t_load.end(t, s_copy);
    slot_of[p] = (int32_t)s;
    part_in_slot[s] = (int32_t)p;
    return LT_OK;
}

// Slot bytes of partition p: the rowptr slice, then the col slice, either
// copied as is, or copied gap-coded into a staging buffer and unpacked from
// there, or (block slots) its block image copied into the slot, where it stays.
// The DMAs run on s_dma; s_copy waits for them (and unpacks), so the load is
// complete when s_copy reaches this point.  A staging buffer is refilled only
// after the unpack of its previous image.  Block slots copy on s_copy alone.
lt_status Engine::copy_partition(uint32_t p, uint8_t *base) {
    const uint32_t nb = part.node_off[p], ne = part.node_off[p + 1];
    const uint64_t eb = part.edge_off[p], ee = part.edge_off[p + 1];
    const size_t rbytes = ((size_t)ne - nb + 1) * 8, cbytes = (size_t)(ee - eb) * 4;
    const bool packed = pk && pk->size[p];
    uint32_t k = 0;
    if (packed && slot_packed) {  // the block image is the slot's column part
        if (rp_resident) {
            LT_HIP(hipMemcpyAsync(base, d_rowptr_all + nb, rbytes, hipMemcpyDeviceToDevice, s_copy));
        } else {
            LT_HIP(hipMemcpyAsync(base, g->rowptr + nb, rbytes, hipMemcpyHostToDevice, s_copy));
            st.h2d_bytes_explicit += rbytes;
            st.h2d_bytes_unpacked += rbytes;
        }
        LT_HIP(hipMemcpyAsync(base + rbytes, pk->host + pk->off[p], pk->size[p], hipMemcpyHostToDevice, s_copy));
        st.h2d_bytes_unpacked += cbytes;
        st.h2d_bytes_explicit += pk->size[p];
        st.packed_image_bytes += pk->size[p];
        st.packed_col_bytes += cbytes;
        return LT_OK;
    }
    if (packed) {
        k = stage_next;
        stage_next ^= 1u;
        if (stage_used[k] && s_dma != s_copy) LT_HIP(hipStreamWaitEvent(s_dma, stage_free[k], 0));
    }
    if (rp_resident) {
        LT_HIP(hipMemcpyAsync(base, d_rowptr_all + nb, rbytes, hipMemcpyDeviceToDevice, s_dma));
    } else {
        LT_HIP(hipMemcpyAsync(base, g->rowptr + nb, rbytes, hipMemcpyHostToDevice, s_dma));
        st.h2d_bytes_explicit += rbytes;
        st.h2d_bytes_unpacked += rbytes;
    }
    st.h2d_bytes_unpacked += cbytes;
    if (packed) {
        const uint64_t n = ne - nb, nch = (ee - eb + 63) / 64;
        const uint64_t nwords = (pk->size[p] - 8 * nch) / 4 - 2;
        const uint64_t waves = (nch + kUnpackChunks - 1) / kUnpackChunks;
        LT_HIP(hipMemcpyAsync(d_stage[k], pk->host + pk->off[p], pk->size[p], hipMemcpyHostToDevice, s_dma));
        if (s_dma != s_copy) {
            LT_HIP(hipEventRecord(ev_dma, s_dma));
            LT_HIP(hipStreamWaitEvent(s_copy, ev_dma, 0));
        }
        unpack_kernel<<<(unsigned)((waves + 3) / 4), 256, 0, s_copy>>>(
            d_stage[k], (uint32_t)n, ee - eb, nwords, pk->vbits, (const uint64_t *)base, (uint32_t *)(base + rbytes));
        LT_HIP(hipGetLastError());
        LT_HIP(hipEventRecord(stage_free[k], s_copy));
        stage_used[k] = true;
        st.h2d_bytes_explicit += pk->size[p];
        st.packed_image_bytes += pk->size[p];
        st.packed_col_bytes += cbytes;
    } else {
        if (cbytes) LT_HIP(hipMemcpyAsync(base + rbytes, g->col + eb, cbytes, hipMemcpyHostToDevice, s_dma));
        st.h2d_bytes_explicit += cbytes;
        if (s_dma != s_copy) {
            LT_HIP(hipEventRecord(ev_dma, s_dma));
            LT_HIP(hipStreamWaitEvent(s_copy, ev_dma, 0));
        }
    }
    return LT_OK;
}

// A partition's slot bytes, built by copy_partition in a scratch buffer and
// copied back; the cache and the statistics are left as they were.
lt_status Engine::fetch_partition(uint32_t p, uint64_t *rowptr_out, uint32_t *col_out) {
    if (slot_packed) {
        set_error("fetch of partition %u: slots keep columns packed (pack_partitions = 2)", p);
        return LT_E_ARG;
    }
    LT_HIP(hipSetDevice(device));
    LT_HIP(hipStreamSynchronize(s_copy));
    if (s_dma != s_copy) LT_HIP(hipStreamSynchronize(s_dma));
    const size_t rn = (size_t)part.node_off[p + 1] - part.node_off[p] + 1;
    const size_t m = part.edge_off[p + 1] - part.edge_off[p];
    uint8_t *buf = nullptr;
    LT_TRY(dmalloc(&buf, rn * 8 + m * 4));
    const lt_stats keep = st;
    lt_status rc = copy_partition(p, buf);
    st = keep;
    hipError_t e = rc == LT_OK ? hipStreamSynchronize(s_copy) : hipSuccess;
    if (rc == LT_OK && e == hipSuccess) e = hipMemcpy(rowptr_out, buf, rn * 8, hipMemcpyDeviceToHost);
    if (rc == LT_OK && e == hipSuccess && m) e = hipMemcpy(col_out, buf + rn * 8, m * 4, hipMemcpyDeviceToHost);
    dfree(buf);
    if (rc != LT_OK) return rc;
    if (e != hipSuccess) {
        set_error("fetch of partition %u: %s", p, hipGetErrorString(e));
        return LT_E_HIP;
    }
    return LT_OK;
}

// Walk launch sized to the slice: 1 walker per thread while the slice is
// small enough that more walkers per thread would leave SIMDs without waves
// (a thread walks its walkers one after another), 2 or 4 for large slices,
// where the per-workgroup append cost (histogram, scan, one atomic per
// destination queue) is spread over more walkers.
template <int APP>
static void launch_walk_app(const WalkArgs &a, int wpt, int grid, size_t lds, hipStream_t s) {
    if (wpt == 4) walk_kernel<APP, 4><<<grid, kBlock, lds, s>>>(a);
    else if (wpt == 2) walk_kernel<APP, 2><<<grid, kBlock, lds, s>>>(a);
    else walk_kernel<APP, 1><<<grid, kBlock, lds, s>>>(a);
}

lt_status Engine::launch_walk(const WalkArgs &a, uint64_t walkers, hipStream_t s) {
    const uint64_t fill = (uint64_t)num_cu * 4 * 6 * 64;  // threads for ~6 waves per SIMD
    const int wpt = cfg.walkers_per_thread > 0 ? cfg.walkers_per_thread
                    : walkers >= 4 * fill         ? 4
                    : walkers >= 2 * fill         ? 2
                                                  : 1;
    const uint64_t bw = (uint64_t)kBlock * wpt;
    const int grid =
        (int)std::max<uint64_t>(1, std::min<uint64_t>((walkers + bw - 1) / bw, (uint64_t)walk_grid[app_kind][wpt]));
    switch (app_kind) {
        case APP_PAGERANK: launch_walk_app<APP_PAGERANK>(a, wpt, grid, lds_bytes(wpt), s); break;
        case APP_PPR: launch_walk_app<APP_PPR>(a, wpt, grid, lds_bytes(wpt), s); break;
        default: launch_walk_app<APP_GENERIC>(a, wpt, grid, lds_bytes(wpt), s); break;
    }
    LT_HIP(hipGetLastError());
    meta_dirty = true;
    return LT_OK;
}

// Consume the whole queue of partition p in slices of kSliceChunks chunks:
// walk kernel, then fix-up kernel per slice (processInMemory with all
// batches, src/WalkManager.h:375-388).
lt_status Engine::process(uint32_t p, bool zerocopy, uint64_t walkers) {
    if (spill_on) walkers = dcounts[p];  // counts[p] also holds the walkers spilled to host
    const uint64_t host = spill_on ? hcount[p] : 0;
    if (walkers == 0 && host == 0) return LT_OK;
    WalkArgs a{};
    a.append_mode = cfg.append_mode;
    a.app = app;
    a.q = q;
    a.pt.node_off = d_node_off;
    a.pt.edge_off = d_edge_off;
    a.pt.lut = d_lut;
    a.pt.rowptr_tab = d_rowptr_tab;
    a.pt.col_tab = d_col_tab;
    a.pt.words_tab = d_words_tab;
    a.pt.zc_part = d_zc_part;
    a.pt.lut_shift = lut_shift;
    a.pt.P = P;
    a.src = p;
    a.zerocopy = zerocopy ? 1 : 0;
    a.span_resident = cfg.span == LT_SPAN_RESIDENT ? 1 : 0;
    a.stats = d_stats;
    a.ss = sched_state(false);
    if (zerocopy) {
        a.rowptr = zc_rowptr + part.node_off[p];
        a.col = zc_col + part.edge_off[p];
    } else {
        const uint8_t *base = d_slots + (size_t)slot_of[p] * slot_sz;
        a.rowptr = (const uint64_t *)base;
        a.col = (const uint32_t *)(base + ((uint64_t)part.node_off[p + 1] - part.node_off[p] + 1) * 8);
        a.words = slot_words(p, a.col);
    }
    EventTimer &phase = zerocopy ? t_zc : t_compute;
    hipEvent_t tp = phase.begin(s_compute);
    const uint64_t nch = (walkers + kChunk - 1) / kChunk;
    for (uint64_t lo = 0; lo < nch; lo += slice) {
        const uint64_t hi = std::min<uint64_t>(nch, lo + slice);
        const bool last = hi == nch;
        a.chunk_lo = (uint32_t)lo;
        a.chunk_hi = last ? 0xFFFFFFFFu : (uint32_t)hi;
        const uint64_t sw = std::min<uint64_t>(walkers, hi * kChunk) - lo * kChunk;
        EventTimer &kt = zerocopy ? t_kzc : t_kernel;
        hipEvent_t tk = kt.begin(s_compute);
        LT_TRY(launch_walk(a, sw, s_compute));
        kt.end(tk, s_compute);
        st.zerocopy_launches += zerocopy ? 1 : 0;
        fixup_kernel<<<1, kBlock, 0, s_compute>>>(q, p, (uint32_t)lo, last ? 0xFFFFFFFFu : (uint32_t)hi, last ? 1 : 0,
                                                  nullptr);
        LT_HIP(hipGetLastError());
        meta_dirty = false;
        st.kernel_launches++;
    }
    if (host) LT_TRY(process_host(p, a));
    phase.end(tp, s_compute);
    return LT_OK;
}

// ---- host walker spill -------------------------------------------------------
lt_status Engine::host_chunk(uint32_t &id) {
    if (hfree.empty()) {
        Walker *b = nullptr, *bd = nullptr;
        LT_HIP(hipHostMalloc((void **)&b, sizeof(Walker) * kChunk * kHostBlockChunks,
                             hipHostMallocMapped | hipHostMallocCoherent));
        hblocks.push_back(b);
        LT_HIP(hipHostGetDevicePointer((void **)&bd, b, 0));
        hblocks_dev.push_back(bd);
        const uint32_t base = (uint32_t)(hblocks.size() - 1) * kHostBlockChunks;
        for (uint32_t k = kHostBlockChunks; k-- > 0;) hfree.push_back(base + k);
    }
    id = hfree.back();
    hfree.pop_back();
    return LT_OK;
}

void Engine::host_reset() {
    for (uint32_t p = 0; p < P; p++) {
        for (uint32_t c : hq[p]) hfree.push_back(c);
        hq[p].clear();
        hhead[p] = hcount[p] = 0;
    }
}

// getEvictedPartition (src/Scheduler.h:149-166): among the partitions other
// than `except` holding at least one full chunk on the device, the one with the
// fewest walkers in all; else any partition with walkers on the device.
int32_t Engine::spill_victim(uint32_t except) const {
    int32_t full = -1, any = -1;
    for (uint32_t i = 0; i < P; i++) {
        if (i == except || dcounts[i] == 0) continue;
        if (dcounts[i] >= kChunk) {
            if (full < 0 || counts[i] < counts[full]) full = (int32_t)i;
        } else if (any < 0 || counts[i] < counts[any]) {
            any = (int32_t)i;
        }
    }
    return full >= 0 ? full : any;
}

// evictWalker (src/Scheduler.h:53-58) for a set of partitions at once: their
// whole device queues go behind their host queues (one gather launch writing
// to mapped host memory), then their chunks return to the pool.  Called with
// the compute stream idle and dcounts fresh.
lt_status Engine::spill_round(const std::vector<uint32_t> &victims) {
    if (victims.empty()) return LT_OK;
    const uint32_t nv = (uint32_t)victims.size();
    std::vector<uint64_t> prefix(nv + 1, 0);
    std::vector<uint32_t> cprefix(nv + 1, 0), hstart(nv), hbase(nv);
    std::vector<Walker *> hptr;
    for (uint32_t k = 0; k < nv; k++) {
        const uint32_t v = victims[k];
        const uint64_t n = dcounts[v], pos = hhead[v] + hcount[v];
        while (((pos + n - 1) >> kChunkLog2) >= hq[v].size()) {
            uint32_t id;
            LT_TRY(host_chunk(id));
            hq[v].push_back(id);
        }
        hstart[k] = (uint32_t)(pos & (kChunk - 1));
        hbase[k] = (uint32_t)hptr.size();
        for (uint64_t c = pos >> kChunkLog2; c <= (pos + n - 1) >> kChunkLog2; c++)
            hptr.push_back(host_chunk_ptr(hq[v][c], true));
        prefix[k + 1] = prefix[k] + n;
        cprefix[k + 1] = cprefix[k] + (uint32_t)((n + kChunk - 1) / kChunk);
    }
    // plan: part | prefix | cprefix | hstart | hbase | hptr, 8-byte aligned pieces
    const size_t o_part = 0, o_prefix = (4 * (size_t)nv + 7) & ~(size_t)7, o_cprefix = o_prefix + 8 * ((size_t)nv + 1);
    const size_t o_hstart = o_cprefix + 4 * ((size_t)nv + 1), o_hbase = o_hstart + 4 * (size_t)nv;
    const size_t o_hptr = (o_hbase + 4 * (size_t)nv + 7) & ~(size_t)7, bytes = o_hptr + 8 * hptr.size();
    if (plan_cap < bytes) {
        dfree(d_plan);
        if (h_plan) (void)hipHostFree(h_plan);
        h_plan = nullptr;
        LT_TRY(dmalloc(&d_plan, bytes));
        LT_HIP(hipHostMalloc((void **)&h_plan, bytes, hipHostMallocDefault));
        plan_cap = bytes;
    }
    memcpy(h_plan + o_part, victims.data(), 4 * (size_t)nv);
    memcpy(h_plan + o_prefix, prefix.data(), 8 * ((size_t)nv + 1));
    memcpy(h_plan + o_cprefix, cprefix.data(), 4 * ((size_t)nv + 1));
    memcpy(h_plan + o_hstart, hstart.data(), 4 * (size_t)nv);
    memcpy(h_plan + o_hbase, hbase.data(), 4 * (size_t)nv);
    memcpy(h_plan + o_hptr, hptr.data(), 8 * hptr.size());
    hipEvent_t te = t_wevict.begin(s_compute);
    LT_HIP(hipMemcpyAsync(d_plan, h_plan, bytes, hipMemcpyHostToDevice, s_compute));
    SpillArgs sa{};
    sa.q = q;
    sa.nv = nv;
    sa.part = (const uint32_t *)(d_plan + o_part);
    sa.prefix = (const uint64_t *)(d_plan + o_prefix);
    sa.cprefix = (const uint32_t *)(d_plan + o_cprefix);
    sa.hstart = (const uint32_t *)(d_plan + o_hstart);
    sa.hbase = (const uint32_t *)(d_plan + o_hbase);
    sa.hptr = (Walker *const *)(d_plan + o_hptr);
    const int grid = (int)std::max<uint64_t>(1, std::min<uint64_t>((prefix[nv] + 255) / 256, (uint64_t)num_cu * 8));
    spill_gather_kernel<<<grid, 256, 0, s_compute>>>(sa);
    LT_HIP(hipGetLastError());
    release_queues_kernel<<<1, kBlock, 0, s_compute>>>(sa);
    LT_HIP(hipGetLastError());
    meta_dirty = false;
    t_wevict.end(te, s_compute);
    LT_HIP(spin_stream(s_compute));  // the plan and host chunks are final before they are reused
    for (uint32_t k = 0; k < nv; k++) {
        hcount[victims[k]] += dcounts[victims[k]];
        dcounts[victims[k]] = 0;
    }
    st.walker_evicting_calls += nv;
    return LT_OK;
}

// handleWalkerPoolOverflow (src/Scheduler.h:59-82): make room for `want`
// walkers (at least half the capacity when want is larger) by spilling device
// queues, never `except`'s.  Once spilling is needed it goes down to half the
// capacity in one round, so that the reloads that follow do not spill again
// one queue at a time.
lt_status Engine::ensure_room(uint64_t want, uint32_t except, uint64_t &room) {
    const uint64_t target = std::max<uint64_t>(1, std::min<uint64_t>(want, dev_cap / 2));
    LT_TRY(read_meta());
    uint64_t live = 0;
    for (uint32_t i = 0; i < P; i++) live += dcounts[i];
    if (live + target > dev_cap) {
        const uint64_t goal = std::min<uint64_t>(dev_cap - target, dev_cap / 2);
        std::vector<uint64_t> keep(dcounts);  // spill_victim reads dcounts: mark victims by zeroing
        std::vector<uint32_t> victims;
        while (live > goal) {
            const int32_t v = spill_victim(except);
            if (v < 0) break;
            live -= dcounts[v];
            victims.push_back((uint32_t)v);
            dcounts[v] = 0;
        }
        dcounts = keep;
        LT_TRY(spill_round(victims));
        LT_TRY(read_meta());
        live = 0;
        for (uint32_t i = 0; i < P; i++) live += dcounts[i];
        if (live + target > dev_cap) {
            set_error("no walkers left to spill (%llu on the device, capacity %llu)", (unsigned long long)live,
                      (unsigned long long)dev_cap);
            return LT_E_INTERNAL;
        }
    }
    room = std::min<uint64_t>(want, dev_cap - live);
    return LT_OK;
}

// walkerLoadingPhase + computeOffMemory (src/Scheduler.h:215-233,
// src/WalkManager.h:389-421): walk p's spilled walkers in launches of at most
// kMaxHostPieces pieces, read from host memory by the kernel, after making
// room on the device for the walkers they may append.
lt_status Engine::process_host(uint32_t p, const WalkArgs &a) {
    while (hcount[p] > 0) {
        uint64_t room;
        LT_TRY(ensure_room(std::min<uint64_t>(hcount[p], (uint64_t)kMaxHostPieces * kChunk), p, room));
        WalkArgs b = a;
        b.nhost = 0;
        uint64_t pos = hhead[p], total = 0;
        while (b.nhost < kMaxHostPieces && total < room) {
            const uint64_t off = pos & (kChunk - 1);
            const uint64_t len = std::min<uint64_t>(kChunk - off, room - total);
            b.hptr[b.nhost] = host_chunk_ptr(hq[p][pos >> kChunkLog2], true) + off;
            b.hn[b.nhost] = (uint32_t)len;
            b.nhost++;
            pos += len;
            total += len;
        }
        hipEvent_t tl = t_wload.begin(s_compute);
        hipEvent_t tk = t_kernel.begin(s_compute);
        LT_TRY(launch_walk(b, total, s_compute));
        t_kernel.end(tk, s_compute);
        fixup_kernel<<<1, kBlock, 0, s_compute>>>(q, kEmpty, 0, 0, 0, nullptr);
        LT_HIP(hipGetLastError());
        meta_dirty = false;
        t_wload.end(tl, s_compute);
        LT_HIP(spin_stream(s_compute));  // the pieces go back to the free list below
        hhead[p] += total;
        hcount[p] -= total;
        while (!hq[p].empty() && (hhead[p] >= kChunk || hcount[p] == 0)) {
            hfree.push_back(hq[p].front());
            hq[p].pop_front();
            hhead[p] = hhead[p] >= kChunk ? hhead[p] - kChunk : 0;
        }
        if (hcount[p] == 0) hhead[p] = 0;
        counts[p] = dcounts[p] + hcount[p];
        spill_h2d_bytes += total * sizeof(Walker);
        st.walker_loading_calls++;
        st.kernel_launches++;
    }
    return LT_OK;
}

// Walk the queue of c over PCIe; with zc_batch > 1 also the next largest
// queues below the threshold (not resident, not being copied), in one launch.
lt_status Engine::process_zc(uint32_t c, const uint8_t *pending) {
    std::vector<uint32_t> batch{c};
    if (zc_batch > 1 && !spill_on) {  // a batch consumes device queues only
        std::vector<uint32_t> cand;
        for (uint32_t i = 0; i < P; i++)
            if (i != c && !resident[i] && !(pending && pending[i]) && counts[i] > 0 && counts[i] < zc_threshold)
                cand.push_back(i);
        const size_t take = std::min<size_t>(cand.size(), zc_batch - 1);
        std::partial_sort(cand.begin(), cand.begin() + take, cand.end(), [&](uint32_t x, uint32_t y) {
            return counts[x] != counts[y] ? counts[x] > counts[y] : x < y;
        });
        uint64_t chunks = (counts[c] + kChunk - 1) / kChunk;
        for (size_t k = 0; k < take; k++) {
            const uint64_t ch = (counts[cand[k]] + kChunk - 1) / kChunk;
            if (chunks + ch > slice) break;  // the pool holds at most `slice` consumed chunks per launch
            chunks += ch;
            batch.push_back(cand[k]);
        }
    }
    for (uint32_t p : batch) record(p, 1, -1, -1);
    st.zerocopy_calls += batch.size();
    if (batch.size() == 1) return process(c, true, counts[c]);
    return walk_batch(batch, true);
}

// Walk the queue of the resident partition c (its decision already
// recorded); with resident_batch > 1 also the largest other resident queues
// of at least one chunk, in one launch, each recorded as a decision of the
// same kind.  Partitions brought in by a prefetch and walked here count as
// prefetch hits.
lt_status Engine::process_resident(uint32_t c, int kind, uint8_t *prefetched) {
    std::vector<uint32_t> batch{c};
    uint64_t chunks = (counts[c] + kChunk - 1) / kChunk;
    if (res_batch > 1 && !spill_on && chunks < slice) {
        std::vector<uint32_t> cand;
        for (uint32_t i = 0; i < P; i++)
            if (i != c && resident[i] && counts[i] >= kChunk) cand.push_back(i);
        const size_t take = std::min<size_t>(cand.size(), res_batch - 1);
        std::partial_sort(cand.begin(), cand.begin() + take, cand.end(), [&](uint32_t x, uint32_t y) {
            return counts[x] != counts[y] ? counts[x] > counts[y] : x < y;
        });
        for (size_t k = 0; k < take; k++) {
            const uint64_t ch = (counts[cand[k]] + kChunk - 1) / kChunk;
            if (chunks + ch > slice) break;  // the pool holds at most `slice` consumed chunks per launch
            chunks += ch;
            batch.push_back(cand[k]);
        }
    }
    if (batch.size() == 1) return process(c, false, counts[c]);
    for (size_t k = 1; k < batch.size(); k++) {
        record(batch[k], kind, -1, -1);
        if (prefetched && prefetched[batch[k]]) {
            st.prefetch_hits++;
            prefetched[batch[k]] = 0;
        }
    }
    st.resident_batched += batch.size() - 1;
    return walk_batch(batch, false);
}

// One launch over the whole queues of `parts` (span resident).  Resident
// partitions are in the partition table already; with zerocopy, the
// partitions are entered in it as views of the pinned host graph for the
// launch (and flagged in zc_part for the PCIe byte count).  Either way a
// walker that moves between them keeps walking, and no queue being consumed
// is appended to.
lt_status Engine::walk_batch(const std::vector<uint32_t> &parts, bool zerocopy) {
    WalkArgs a{};
    a.append_mode = cfg.append_mode;
    a.app = app;
    a.q = q;
    a.pt.node_off = d_node_off;
    a.pt.edge_off = d_edge_off;
    a.pt.lut = d_lut;
    a.pt.rowptr_tab = d_rowptr_tab;
    a.pt.col_tab = d_col_tab;
    a.pt.words_tab = d_words_tab;
    a.pt.zc_part = d_zc_part;
    a.pt.lut_shift = lut_shift;
    a.pt.P = P;
    a.span_resident = 1;
    a.stats = d_stats;
    a.ss = sched_state(false);
    ZcBatch b{};
    b.n = (uint32_t)parts.size();
    a.nsrc = b.n;
    a.src = parts[0];
    a.voff[0] = 0;
    uint64_t total = 0;
    for (uint32_t k = 0; k < b.n; k++) {
        const uint32_t p = parts[k];
        b.part[k] = p;
        b.rowptr[k] = zc_rowptr + part.node_off[p];
        b.col[k] = zc_col + part.edge_off[p];
        a.srcs[k] = p;
        a.voff[k + 1] = a.voff[k] + (counts[p] + kChunk - 1) / kChunk * kChunk;
        total += counts[p];
    }
    EventTimer &phase = zerocopy ? t_zc : t_compute;
    EventTimer &kt = zerocopy ? t_kzc : t_kernel;
    hipEvent_t tp = phase.begin(s_compute);
    if (zerocopy) {
        set_zc_tables_kernel<<<1, 64, 0, s_compute>>>(d_rowptr_tab, d_col_tab, d_zc_part, b, 1);
        LT_HIP(hipGetLastError());
    }
    hipEvent_t tk = kt.begin(s_compute);
    LT_TRY(launch_walk(a, total, s_compute));
    kt.end(tk, s_compute);
    st.zerocopy_launches += zerocopy ? 1 : 0;
    for (uint32_t p : parts) fixup_kernel<<<1, kBlock, 0, s_compute>>>(q, p, 0, 0xFFFFFFFFu, 1, nullptr);
    LT_HIP(hipGetLastError());
    meta_dirty = false;
    if (zerocopy) {
        set_zc_tables_kernel<<<1, 64, 0, s_compute>>>(d_rowptr_tab, d_col_tab, d_zc_part, b, 0);
        LT_HIP(hipGetLastError());
    }
    st.kernel_launches++;
    phase.end(tp, s_compute);
    return LT_OK;
}

// One decision: with tracing, the choice and the state it was taken on; then
// a walk from HBM (every kind but 1, zero-copy) advances p's recency.
void Engine::record(uint32_t p, int kind, int32_t loaded, int32_t evicted) {
    if (cfg.record_trace) {
        trace.push_back(lt_trace_rec{(int32_t)p, kind, loaded, evicted, counts[p]});
        for (uint32_t i = 0; i < P; i++) trace_counts.push_back((uint32_t)counts[i]);
        trace_resident.insert(trace_resident.end(), resident.begin(), resident.end());
        for (uint32_t i = 0; i < P; i++) trace_age.push_back(age_of(i));
        int32_t rec[kRecent];
        recent_loads(rec);
        trace_recent.insert(trace_recent.end(), rec, rec + kRecent);
    }
    if (kind != 1) last_use[p] = ++use_clock;
}

uint32_t Engine::age_of(uint32_t p) const {
    return (uint32_t)std::min<uint64_t>(use_clock - last_use[p], 0xFFFFFFFFu);
}

void Engine::recent_loads(int32_t *out) const {
    const auto &h = sched.history;
    for (int k = 0; k < kRecent; k++) out[k] = k < (int)h.size() ? h[h.size() - 1 - k] : -1;
}

// Scores of every partition for the decision that chose `cur`.
lt_status Engine::infer_scores(int32_t cur) {
    scores_valid = false;
    if (!scorer || !scorer_trained(scorer)) return LT_OK;
    LT_TRY(infer_launch(cur));
    return infer_collect();
}

// Launch the inference of the decision that chose `cur` without waiting for
// it (the state is copied at once); infer_collect makes its scores current.
lt_status Engine::infer_launch(int32_t cur) {
    LT_TRY(infer_collect());  // one inference in flight
    if (!scorer || !scorer_trained(scorer)) return LT_OK;
    std::vector<uint32_t> c32(P), a32(P);
    for (uint32_t i = 0; i < P; i++) {
        c32[i] = (uint32_t)std::min<uint64_t>(counts[i], 0xFFFFFFFFu);
        a32[i] = age_of(i);
    }
    DecisionState d;
    d.counts = c32.data();
    d.resident = resident.data();
    d.age = a32.data();
    recent_loads(d.recent);
    d.cur = cur;
    d.w_total = (float)w_run;
    LT_TRY(scorer_infer_launch(scorer, d, nullptr));
    infer_pending = true;
    pending_clock = use_clock;
    return LT_OK;
}

lt_status Engine::infer_collect() {
    if (!infer_pending) return LT_OK;
    infer_pending = false;
    scores_valid = false;
    // another caller of the scorer (after a run that ended early) waited for it
    if (!scorer || !scorer_infer_pending(scorer)) return LT_OK;
    scores.resize((size_t)P * kOut);
    LT_TRY(scorer_infer_wait(scorer, scores.data()));
    scores_clock = pending_clock;
    bool ok = true;
    for (float v : scores) ok = ok && std::isfinite(v);
    scores_valid = ok;
    return LT_OK;
}

// Victim among the resident partitions other than a and b: those with the
// fewest queued walkers; of them the evict_shortlist least recently walked
// (all of them with 0, the default); of those the one with the lowest keep
// score (ties: lowest id).  A shortlist bounds the network's mistakes on
// states unlike its training traces (a network trained on Belady's choices
// for the plain scheduler's states can rank a partition walked moments ago
// lowest on the states its own choices lead to).  Without valid scores, the
// plain rule.
int32_t Engine::learned_victim(int32_t a, int32_t b) {
    if (scores_valid) {
        uint64_t fewest = UINT64_MAX;
        for (uint32_t i = 0; i < P; i++)
            if (resident[i] && (int32_t)i != a && (int32_t)i != b) fewest = std::min(fewest, counts[i]);
        if (fewest != UINT64_MAX) {
            std::vector<std::pair<uint32_t, int32_t>> cand;  // (age, id)
            for (uint32_t i = 0; i < P; i++)
                if (resident[i] && (int32_t)i != a && (int32_t)i != b && counts[i] == fewest)
                    cand.push_back({age_of(i), (int32_t)i});
            const size_t k = cfg.evict_shortlist > 0 ? (size_t)cfg.evict_shortlist : cand.size();
            if (k < cand.size()) {  // oldest first, ties by id
                std::partial_sort(cand.begin(), cand.begin() + k, cand.end(),
                                  [](const std::pair<uint32_t, int32_t> &x, const std::pair<uint32_t, int32_t> &y) {
                                      return x.first > y.first || (x.first == y.first && x.second < y.second);
                                  });
                cand.resize(k);
            }
            int32_t v = -1;
            float keep = 0.f;
            for (const auto &e : cand) {
                const float kv = scores[(size_t)e.second * kOut + 1];
                if (v < 0 || kv < keep || (kv == keep && e.second < v)) {
                    keep = kv;
                    v = e.second;
                }
            }
            sched.learned_decisions++;
            return v;
        }
    }
    sched.fallback_decisions++;
    return lt_sched_evict_min(P, counts.data(), resident.data(), a);
}

// GPUScheduler::run (src/Scheduler.h:311-330) without host walker spill.
lt_status Engine::run_partition(uint32_t p, int32_t &prev_loaded) {
    st.iterations++;
    const uint64_t n = counts[p];
    // shouldZerocopy (src/Scheduler.h:243-252): counts were read this iteration.
    if (!resident[p] && n < zc_threshold) return process_zc(p, nullptr);
    st.explicit_iterations++;
    // graphLoadingPhase (src/Scheduler.h:173-213)
    if (!resident[p]) {
        int32_t evicted = -1;
        if (free_slots.empty()) {
            const int32_t victim = sched.evict_choice(counts.data(), resident.data(), prev_loaded);
            if (victim < 0) {
                set_error("no partition to evict");
                return LT_E_INTERNAL;
            }
            LT_TRY(evict((uint32_t)victim));
            evicted = victim;
        }
        record(p, 3, (int32_t)p, evicted);
        LT_TRY(load_async(p));
        st.graph_loading_calls++;
        sched.on_load(p);
        if (cfg.pipeline) {
            // compute resident partitions while the copy is in flight
            while (hipStreamQuery(s_copy) == hipErrorNotReady) {
                LT_TRY(read_meta());
                const int32_t c = sched.computed_choice(counts.data(), resident.data(), (int32_t)p, kRefPage);
                if (c < 0) break;
                st.computing_calls++;
                record((uint32_t)c, 2, -1, -1);
                LT_TRY(process_resident((uint32_t)c, 2, nullptr));
                LT_HIP(hipStreamSynchronize(s_compute));
            }
            // the pipelined work changed the queues; p's count is re-read below
            LT_TRY(read_meta());
        }
        LT_HIP(hipStreamSynchronize(s_copy));
        resident[p] = 1;
        LT_TRY(set_table(p, true));
    } else {
        record(p, 0, -1, -1);
    }
    prev_loaded = (int32_t)p;
    st.computing_calls++;
    return process_resident(p, 0, nullptr);
}

// Scheduler::start (src/Scheduler.h:425-438).
lt_status Engine::scheduler_loop() {
    int32_t prev = -1, prev_loaded = -1;
    bool sync = false;
    sched.begin_run();
    while (true) {
        LT_TRY(read_meta());
        const int32_t c = sched.next(counts.data(), prev, sync);
        if (c == -2) return LT_E_INTERNAL;
        if (c < 0) {
            if (sync) break;
            sync = true;
            continue;
        }
        sync = false;
        prev = c;
        LT_TRY(run_partition((uint32_t)c, prev_loaded));
    }
    LT_HIP(hipDeviceSynchronize());
    return LT_OK;
}

// Learned scheduling.  Decisions follow the plain scheduler (choice, zero-copy
// rule, pipelined compute during a load), with two changes: while the chosen
// partition is walked, the partitions that will need host data next (the
// scorer's best guesses, or with prefetch_plain the largest such queues) are
// copied into slots ahead of time (prefetch), and a victim is the resident
// partition with the fewest queued walkers, ties broken by the lowest keep
// score: trained on Belady's next-use labels, it names the partition needed
// farthest ahead (the plain rule breaks ties by partition id).  Evictions happen only while the compute stream is idle, and
// a prefetched slot joins the resident table only after its copy completed.
lt_status Engine::scheduler_loop_learned() {
    int32_t prev = -1, prev_loaded = -1;
    bool sync = false;
    struct Inflight {
        int32_t part;
        hipEvent_t done;
    };
    std::vector<Inflight> inflight;         // prefetch copies in copy-stream order
    std::vector<uint8_t> prefetched(P, 0);  // resident because of a prefetch, not yet picked
    std::vector<uint8_t> pending(P, 0);     // being copied by a prefetch
    const uint32_t depth = (uint32_t)std::max(0, cfg.prefetch_depth);
    // Prefetch targets by the plain rule (prefetch_plain 1 and 2): the
    // largest non-resident queues.  With prefetch_plain 1 the victims follow
    // the plain rule too and the scorer is not used (an ablation that
    // isolates what the network adds); with 2 the scorer picks the victims.
    const bool plain_victims = cfg.prefetch_plain == 1;
    const bool plain_targets = cfg.prefetch_plain != 0;
    LT_TRY(infer_collect());
    scores_valid = false;
    sched.begin_run();
    // make completed prefetches resident, in order; with `upto`, wait for it
    auto retire = [&](int32_t upto) -> lt_status {
        while (!inflight.empty()) {
            Inflight f = inflight.front();
            const bool must = upto >= 0;
            if (must) LT_HIP(spin_event(f.done));
            else if (hipEventQuery(f.done) != hipSuccess) break;
            inflight.erase(inflight.begin());
            pool_events.push_back(f.done);
            pending[f.part] = 0;
            resident[f.part] = 1;
            LT_TRY(set_table((uint32_t)f.part, true));
            if (f.part == upto) break;
        }
        return LT_OK;
    };
    // reference pipelining: walk resident partitions while the copy engine works
    auto pipeline_while_copying = [&](int32_t except) -> lt_status {
        if (!cfg.pipeline) return LT_OK;
        while (hipStreamQuery(s_copy) == hipErrorNotReady) {
            LT_TRY(read_meta());
            const int32_t p2 = sched.computed_choice(counts.data(), resident.data(), except, kRefPage);
            if (p2 < 0) break;
            st.computing_calls++;
            record((uint32_t)p2, 2, -1, -1);
            LT_TRY(process_resident((uint32_t)p2, 2, prefetched.data()));
            LT_HIP(hipStreamSynchronize(s_compute));
        }
        return read_meta();
    };
    while (true) {
        LT_TRY(read_meta());
        LT_TRY(retire(-1));
        LT_TRY(infer_collect());  // the previous decision's scores, computed while its walk ran
        int32_t c = sched.next(counts.data(), prev, sync);
        if (c < 0) {
            if (sync) break;
            sync = true;
            continue;
        }
        // Stick to a prefetched partition (copied or in flight) whose queue is
        // close to the largest: its data is already on the way.
        if (cfg.prefetch_stick_pct > 0 && !resident[c] && !pending[c]) {
            int32_t s2 = -1;
            for (uint32_t i = 0; i < P; i++)
                if ((pending[i] || prefetched[i]) && counts[i] >= zc_threshold && (s2 < 0 || counts[i] > counts[s2]))
                    s2 = (int32_t)i;
            if (s2 >= 0 && counts[s2] * 100 >= counts[c] * (uint64_t)cfg.prefetch_stick_pct) c = s2;
        }
        // The scorer's next partition (learned_next): when c still needs host
        // data, walk instead the highest "next" score among the partitions
        // whose queue is within learned_next % of c's (the greedy rule's
        // exclusion of the previous partition kept).  Scores are the last
        // inference's, normally the previous decision's.
        if (cfg.learned_next > 0 && !plain_victims && !resident[c] && !pending[c] && scorer &&
            scorer_trained(scorer)) {
            if (!scores_recent()) LT_TRY(infer_scores(prev));
            if (scores_valid) {
                int32_t s2 = -1;
                float best = 0.f;
                for (uint32_t i = 0; i < P; i++) {
                    if (counts[i] == 0 || counts[i] * 100 < counts[c] * (uint64_t)cfg.learned_next) continue;
                    if ((int32_t)i == prev && !sync) continue;
                    const float v = scores[(size_t)i * kOut];
                    if (s2 < 0 || v > best) {
                        best = v;
                        s2 = (int32_t)i;
                    }
                }
                if (s2 >= 0 && s2 != c) {
                    st.next_overrides++;
                    c = s2;
                }
            }
        }
        sync = false;
        prev = c;
        st.iterations++;
        bool zc = false;
        if (!resident[c] && !pending[c] && counts[c] < zc_threshold) {
            zc = true;
        } else {
            st.explicit_iterations++;
            if (pending[c]) {
                if (hipEventQuery(inflight.front().done) != hipSuccess && inflight.front().part == c)
                    LT_TRY(pipeline_while_copying(c));
                LT_TRY(retire(c));
            }
            if (resident[c]) {
                if (prefetched[c]) st.prefetch_hits++;
                prefetched[c] = 0;
                record((uint32_t)c, 0, -1, -1);
            } else {
                LT_TRY(retire(inflight.empty() ? -1 : inflight.back().part));  // the copy engine is in-order
                int32_t evicted = -1;
                if (free_slots.empty()) {
                    if (!scores_recent()) LT_TRY(infer_scores(c));
                    evicted = learned_victim(c, prev_loaded);
                    if (evicted < 0) {
                        set_error("no partition to evict");
                        return LT_E_INTERNAL;
                    }
                    LT_TRY(evict((uint32_t)evicted));
                    prefetched[evicted] = 0;
                }
                record((uint32_t)c, 3, c, evicted);
                LT_TRY(load_async((uint32_t)c));
                st.graph_loading_calls++;
                sched.on_load((uint32_t)c);
                LT_TRY(pipeline_while_copying(c));
                LT_HIP(hipStreamSynchronize(s_copy));
                resident[c] = 1;
                LT_TRY(set_table((uint32_t)c, true));
            }
            prev_loaded = c;
        }
        // Prefetch the predicted next loads while c is walked.  Slots are
        // freed before the walk is launched (the walk then never reads them),
        // with the scores of the previous decision; the targets are chosen
        // with the scores of this decision, computed while the walk runs.
        size_t want = (plain_victims || (scorer && scorer_trained(scorer))) && inflight.size() < depth
                          ? depth - inflight.size()
                          : 0;
        if (want) {  // no more slots than there are partitions a prefetch may take
            size_t cand = 0;
            for (uint32_t i = 0; i < P && cand < want; i++)
                cand += !resident[i] && !pending[i] && (int32_t)i != c && counts[i] >= zc_threshold;
            want = cand;
        }
        if (want) {
            // the scores of the last inference (normally the previous
            // decision's, computed while its walk ran): no inference here on
            // the critical path unless they are far out of date
            if (!scores_recent() && !plain_victims) LT_TRY(infer_scores(c));
            bool evicted_any = false;
            while ((scores_valid || plain_victims) && free_slots.size() < want) {
                const int32_t v = learned_victim(c, -1);
                if (v < 0) break;
                LT_TRY(evict((uint32_t)v));
                prefetched[v] = 0;
                evicted_any = true;
            }
            if (evicted_any) {  // copies into the freed slots start after the table updates
                if (!ev_evicted) LT_HIP(hipEventCreateWithFlags(&ev_evicted, hipEventDisableTiming));
                LT_HIP(hipEventRecord(ev_evicted, s_compute));
                LT_HIP(hipStreamWaitEvent(s_copy, ev_evicted, 0));
                if (s_dma != s_copy) LT_HIP(hipStreamWaitEvent(s_dma, ev_evicted, 0));
            }
        }
        st.computing_calls++;
        LT_TRY(zc ? process_zc((uint32_t)c, pending.data()) : process_resident((uint32_t)c, 0, prefetched.data()));
        // this decision's scores, computed while c is walked: the next
        // decision's victims, and (prefetch_plain 0) this decision's targets
        if (!plain_victims) LT_TRY(infer_launch(c));
        if (want && !free_slots.empty()) {
            if (!plain_targets) LT_TRY(infer_collect());
            while ((scores_valid || plain_targets) && inflight.size() < depth && !free_slots.empty()) {
                int32_t q = -1;
                float best = 0.f;
                for (uint32_t i = 0; i < P; i++) {
                    if (resident[i] || pending[i] || (int32_t)i == c || counts[i] < zc_threshold) continue;
                    // the scorer's guess of the next partition needing host data, or
                    // (plain rule) the largest queue among those partitions
                    const float v = plain_targets ? (float)counts[i] : scores[(size_t)i * kOut];
                    if (q < 0 || v > best) {
                        best = v;
                        q = (int32_t)i;
                    }
                }
                if (q < 0) break;
                LT_TRY(load_async((uint32_t)q));
                hipEvent_t ev;
                LT_TRY(take_event(ev));
                LT_HIP(hipEventRecord(ev, s_copy));
                inflight.push_back({q, ev});
                pending[q] = 1;
                st.graph_loading_calls++;
                sched.on_load((uint32_t)q);
                st.prefetch_issued++;
                prefetched[q] = 1;
            }
        }
    }
    LT_TRY(retire(inflight.empty() ? -1 : inflight.back().part));
    LT_TRY(infer_collect());
    LT_HIP(hipDeviceSynchronize());
    return LT_OK;
}

SchedState Engine::sched_state(bool device) const {
    SchedState ss{};
    ss.ds = device ? d_ds : nullptr;
    ss.slot_of = d_slot_of;
    ss.part_in_slot = d_part_in_slot;
    ss.resident = d_resident;
    ss.rowptr_tab = d_rowptr_tab;
    ss.col_tab = d_col_tab;
    ss.slots = d_slots;
    ss.slot_bytes = slot_sz;
    ss.nslots = nslots;
    ss.zc_rowptr = zc_rowptr;
    ss.rp_host = rp_resident ? 0 : 1;
    ss.zc_col = zc_col;
    ss.zc_threshold = zc_threshold;
    return ss;
}

// Device-scheduled loop: per iteration decide_kernel (plain greedy rules in
// one workgroup), load_kernel (partition copy over PCIe by the GPU itself),
// walk_kernel and fixup_kernel, all reading the decision from device memory.
// kGraphIters iterations are captured once into a hipGraph and the graph is
// relaunched until an iteration finds every queue empty; the host reads one
// 40-byte record per launch instead of syncing on every decision.
constexpr int kGraphIters = 32;

lt_status Engine::scheduler_loop_device() {
    // host residency -> device mirrors
    std::vector<int32_t> hs(P), hp(nslots);
    for (uint32_t i = 0; i < P; i++) hs[i] = resident[i] ? slot_of[i] : -1;
    for (uint32_t s2 = 0; s2 < nslots; s2++) hp[s2] = part_in_slot[s2];
    LT_HIP(hipMemcpyAsync(d_slot_of, hs.data(), 4 * P, hipMemcpyHostToDevice, s_compute));
    LT_HIP(hipMemcpyAsync(d_part_in_slot, hp.data(), 4 * nslots, hipMemcpyHostToDevice, s_compute));
    LT_HIP(hipMemcpyAsync(d_resident, resident.data(), P, hipMemcpyHostToDevice, s_compute));
    DevSched zero{};
    zero.choice = 0;
    LT_HIP(hipMemcpyAsync(d_ds, &zero, sizeof(zero), hipMemcpyHostToDevice, s_compute));
    LT_HIP(hipStreamSynchronize(s_compute));

    WalkArgs a{};
    a.append_mode = cfg.append_mode;
    a.app = app;
    a.q = q;
    a.pt.node_off = d_node_off;
    a.pt.edge_off = d_edge_off;
    a.pt.lut = d_lut;
    a.pt.rowptr_tab = d_rowptr_tab;
    a.pt.col_tab = d_col_tab;
    a.pt.words_tab = d_words_tab;
    a.pt.zc_part = d_zc_part;
    a.pt.lut_shift = lut_shift;
    a.pt.P = P;
    a.span_resident = cfg.span == LT_SPAN_RESIDENT ? 1 : 0;
    a.stats = d_stats;
    a.ss = sched_state(true);
    const int load_grid = 4 * num_cu;
    // walkers per lane of the captured walk launches: the queue sizes are only
    // known on the device, so the grid is the full occupancy grid of that WPT
    const int dwpt = cfg.walkers_per_thread > 0 ? cfg.walkers_per_thread : 1;
    BlockImages bi{};
    bi.words_tab = d_words_tab;
    bi.img = d_pk_img;
    bi.off = d_pk_off;
    bi.size = d_pk_size;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    LT_HIP(hipStreamBeginCapture(s_compute, hipStreamCaptureModeThreadLocal));
    for (int it = 0; it < kGraphIters; it++) {
        decide_kernel<<<1, 1024, 0, s_compute>>>(q, a.ss, a.pt, bi);
        load_kernel<<<load_grid, 256, 0, s_compute>>>(a.ss, a.pt, bi);
        switch (app_kind) {
            case APP_PAGERANK:
                launch_walk_app<APP_PAGERANK>(a, dwpt, walk_grid[APP_PAGERANK][dwpt], lds_bytes(dwpt), s_compute);
                break;
            case APP_PPR:
                launch_walk_app<APP_PPR>(a, dwpt, walk_grid[APP_PPR][dwpt], lds_bytes(dwpt), s_compute);
                break;
            default:
                launch_walk_app<APP_GENERIC>(a, dwpt, walk_grid[APP_GENERIC][dwpt], lds_bytes(dwpt), s_compute);
                break;
        }
        fixup_kernel<<<1, kBlock, 0, s_compute>>>(q, kEmpty, 0, 0, 0, d_ds);
    }
    hipError_t ce = hipStreamEndCapture(s_compute, &graph);
    if (ce != hipSuccess) {
        set_error("graph capture failed: %s", hipGetErrorString(ce));
        return LT_E_HIP;
    }
    std::unique_ptr<hipGraph_t, void (*)(hipGraph_t *)> gg(&graph, [](hipGraph_t *x) { (void)hipGraphDestroy(*x); });
    LT_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    std::unique_ptr<hipGraphExec_t, void (*)(hipGraphExec_t *)> ge(&exec,
                                                                    [](hipGraphExec_t *x) { (void)hipGraphExecDestroy(*x); });
    DevSched h{};
    uint64_t launches = 0;
    while (true) {
        LT_HIP(hipGraphLaunch(exec, s_compute));
        launches++;
        LT_HIP(hipMemcpyAsync(&h, d_ds, sizeof(h), hipMemcpyDeviceToHost, s_compute));
        LT_HIP(hipStreamSynchronize(s_compute));
        if (h.choice < 0) break;
        if (launches > (1ull << 24)) {
            set_error("device schedule did not drain the queues");
            return LT_E_INTERNAL;
        }
    }
    LT_TRY(read_meta());
    // device residency -> host mirrors
    LT_HIP(hipMemcpy(hs.data(), d_slot_of, 4 * P, hipMemcpyDeviceToHost));
    LT_HIP(hipMemcpy(hp.data(), d_part_in_slot, 4 * nslots, hipMemcpyDeviceToHost));
    LT_HIP(hipMemcpy(resident.data(), d_resident, P, hipMemcpyDeviceToHost));
    free_slots.clear();
    for (uint32_t i = 0; i < P; i++) slot_of[i] = resident[i] ? hs[i] : -1;
    for (int32_t s2 = (int32_t)nslots - 1; s2 >= 0; s2--) {
        part_in_slot[s2] = hp[s2];
        if (hp[s2] < 0) free_slots.push_back((uint32_t)s2);
    }
    st.iterations += h.iterations;
    st.explicit_iterations += h.iterations - h.zc_calls;
    st.graph_loading_calls += h.loads;
    st.zerocopy_calls += h.zc_calls;
    st.computing_calls += h.iterations;
    st.h2d_bytes_explicit += h.h2d_bytes;
    st.h2d_bytes_unpacked += h.h2d_bytes;
    st.kernel_launches += (uint64_t)launches * kGraphIters;
    return LT_OK;
}

lt_status Engine::run(const RunSpec &rs, uint32_t *counts_out, uint32_t *endpoints_out, uint32_t *moves_out) {
    LT_HIP(hipSetDevice(device));
    memset(&st, 0, sizeof(st));
    st.cache_slots = nslots;
    spill_h2d_bytes = 0;
    st.h2d_bytes_explicit = pending_h2d;  // the rowptr upload, once
    st.h2d_bytes_unpacked = pending_h2d;
    pending_h2d = 0;
    trace.clear();
    trace_counts.clear();
    trace_resident.clear();
    trace_recent.clear();
    trace_age.clear();
    t_load.reset();
    t_compute.reset();
    t_zc.reset();
    t_wload.reset();
    t_wevict.reset();
    t_kernel.reset();
    t_kzc.reset();
    const uint64_t cap = cfg.num_device_walkers ? std::min<uint64_t>(cfg.num_device_walkers, rs.total) : rs.total;
    // More walkers than the device holds: spill to host queues (the
    // reference's -b); the device-scheduled loop runs them in waves instead.
    spill_on = cap < rs.total && !cfg.device_schedule;
    dev_cap = cap;
    const uint64_t wave = spill_on ? rs.total : cap;
    // Per-XCD count copies pay when the walkers of one launch stay in a few
    // partitions (capped cache, PPR around its sources); with every partition
    // resident the visits spread over all of V and one array is faster.
    count_copies = rs.app != APP_GENERIC && (nslots < P || rs.app == APP_PPR) ? kCountCopies : 1;
    LT_TRY(ensure_pool(std::max<uint64_t>(cap, 1)));
    LT_TRY(ensure_outputs(rs, endpoints_out != nullptr, moves_out != nullptr));

    app_kind = rs.app;
    app = AppArgs{};
    app.seed = cfg.seed;
    app.V = V;
    app.length = rs.length;
    app.source = rs.source;
    app.counts = d_counts;
    app.counts_stride = count_copies > 1 ? V : 0;
    app.endpoints = endpoints_out ? d_endpoints : nullptr;
    app.moves_out = moves_out ? d_moves : nullptr;
    app.paths = rs.paths_out ? d_paths : nullptr;
    app.rp_host = rp_resident ? 0 : 1;
    if (rs.app == APP_PAGERANK) app.thresh = prob_threshold(1.0 - (double)0.85f);
    else if (rs.app == APP_PPR) app.thresh = prob_threshold((double)rs.prob);
    if (rs.app == APP_PPR && app.thresh == 0) {
        set_error("ppr stop probability %g rounds to zero", (double)rs.prob);
        return LT_E_ARG;
    }
    if (rs.nq) {
        std::vector<uint64_t> off(rs.nq + 1, 0);
        for (uint32_t i = 0; i < rs.nq; i++) off[i + 1] = off[i] + rs.q_cnt[i];
        LT_HIP(hipMemcpy(d_qoff, off.data(), 8 * (rs.nq + 1), hipMemcpyHostToDevice));
        LT_HIP(hipMemcpy(d_qsrc, rs.q_src, 4 * rs.nq, hipMemcpyHostToDevice));
        app.q_off = d_qoff;
        app.q_src = d_qsrc;
        app.nq = rs.nq;
    }
    LT_HIP(hipMemsetAsync(d_counts, 0, 4 * (size_t)V * count_copies, s_compute));
    LT_HIP(hipMemsetAsync(d_stats, 0, 8 * (size_t)kNumStats * kStatSlots, s_compute));
    if (moves_out) LT_HIP(hipMemsetAsync(d_moves, 0, 4 * rs.total, s_compute));
    if (rs.paths_out)  // columns after an early stop stay 0xFFFFFFFF
        LT_HIP(hipMemsetAsync(d_paths, 0xFF, 4 * (size_t)rs.total * (rs.length + 1), s_compute));

    double running_ms = 0;
    for (uint64_t id0 = 0; id0 < rs.total; id0 += wave) {
        const uint64_t id1 = std::min<uint64_t>(rs.total, id0 + wave);
        // createWalkers: untimed, as in the reference (src/LightTraffic.h:33-44)
        reset_queues_kernel<<<1024, 256, 0, s_compute>>>(q, P);
        LT_HIP(hipGetLastError());
        meta_dirty = true;
        InitArgs ia{};
        ia.app = app;
        ia.q = q;
        ia.pt.node_off = d_node_off;
        ia.pt.edge_off = d_edge_off;
        ia.pt.lut = d_lut;
        ia.pt.rowptr_tab = d_rowptr_tab;
        ia.pt.col_tab = d_col_tab;
        ia.pt.words_tab = d_words_tab;
        ia.pt.zc_part = d_zc_part;
        ia.pt.lut_shift = lut_shift;
        ia.pt.P = P;
        ia.stats = d_stats;
        host_reset();
        for (uint64_t lo = id0; lo < id1;) {  // with spill, in pieces that fit on the device
            uint64_t n = id1 - lo;
            if (spill_on) LT_TRY(ensure_room(n, kEmpty, n));
            ia.id_lo = lo;
            ia.id_hi = lo + n;
            const int grid = (int)std::min<uint64_t>((n + kBlockWalkers - 1) / kBlockWalkers, (uint64_t)max_grid);
            switch (rs.app) {
                case APP_PAGERANK: init_kernel<APP_PAGERANK><<<grid, kBlock, lds_bytes(), s_compute>>>(ia); break;
                case APP_PPR: init_kernel<APP_PPR><<<grid, kBlock, lds_bytes(), s_compute>>>(ia); break;
                default: init_kernel<APP_GENERIC><<<grid, kBlock, lds_bytes(), s_compute>>>(ia); break;
            }
            LT_HIP(hipGetLastError());
            fixup_kernel<<<1, kBlock, 0, s_compute>>>(q, kEmpty, 0, 0, 0, nullptr);
            LT_HIP(hipGetLastError());
            meta_dirty = false;
            lo += n;
        }
        LT_HIP(hipStreamSynchronize(s_compute));
        const auto t0 = std::chrono::steady_clock::now();
        w_run = id1 - id0;
        if (cfg.device_schedule) LT_TRY(scheduler_loop_device());
        else LT_TRY(cfg.policy == LT_SCHED_LEARNED && (scorer || cfg.prefetch_plain) ? scheduler_loop_learned()
                                                                                   : scheduler_loop());
        const auto t1 = std::chrono::steady_clock::now();
        running_ms += std::chrono::duration<double, std::milli>(t1 - t0).count();
    }
    if (count_copies > 1) {  // part of the walk's work, so inside the timed region
        const auto t0 = std::chrono::steady_clock::now();
        sum_count_copies_kernel<<<std::max(1, 8 * num_cu), 256, 0, s_compute>>>(d_counts, V);
        LT_HIP(hipGetLastError());
        LT_HIP(hipStreamSynchronize(s_compute));
        running_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    LT_TRY(read_meta());
    for (uint32_t i = 0; i < P; i++)
        if (counts[i]) {
            set_error("partition %u still holds %llu walkers after the run", i, (unsigned long long)counts[i]);
            return LT_E_INTERNAL;
        }

    // reduce (app/pagerank.cu:71-74): results back to the host, untimed
    if (counts_out) LT_HIP(hipMemcpy(counts_out, d_counts, 4 * (size_t)V, hipMemcpyDeviceToHost));
    if (endpoints_out) LT_HIP(hipMemcpy(endpoints_out, d_endpoints, 4 * rs.total, hipMemcpyDeviceToHost));
    if (moves_out) LT_HIP(hipMemcpy(moves_out, d_moves, 4 * rs.total, hipMemcpyDeviceToHost));
    if (rs.paths_out)
        LT_HIP(hipMemcpy(rs.paths_out, d_paths, 4 * (size_t)rs.total * (rs.length + 1), hipMemcpyDeviceToHost));
    std::vector<unsigned long long> slots_h((size_t)kNumStats * kStatSlots);
    LT_HIP(hipMemcpy(slots_h.data(), d_stats, 8 * slots_h.size(), hipMemcpyDeviceToHost));
    unsigned long long hs[kNumStats] = {};
    for (size_t i = 0; i < slots_h.size(); i++) hs[i % kNumStats] += slots_h[i];
    st.running_ms = running_ms;
    st.moves = hs[ST_MOVES];
    st.teleports = hs[ST_TELEPORTS];
    st.visits = hs[ST_VISITS];
    st.residencies = hs[ST_RESIDENCIES];
    st.zerocopy_moves = hs[ST_ZC_MOVES];
    st.h2d_bytes_zerocopy = hs[ST_ZC_READS] * 64ull + spill_h2d_bytes;
    st.nominal_steps = rs.nominal;
    t_load.collect();
    t_compute.collect();
    t_zc.collect();
    t_kernel.collect();
    t_kzc.collect();
    st.graph_loading_ms = t_load.ms;
    st.computing_ms = t_compute.ms;
    st.zerocopy_ms = t_zc.ms;
    t_wload.collect();
    t_wevict.collect();
    st.walker_loading_ms = t_wload.ms;
    st.walker_evicting_ms = t_wevict.ms;
    st.walk_kernel_ms = t_kernel.ms + t_kzc.ms;
    st.walk_kernel_zc_ms = t_kzc.ms;
    st.learned_decisions = sched.learned_decisions;
    st.fallback_decisions = sched.fallback_decisions;
    return LT_OK;
}

}  // namespace lt

struct lt_engine : lt::Engine {};

using namespace lt;

extern "C" {

lt_status lt_engine_create(const lt_graph *g, const lt_partition *p, const lt_config *cfg, lt_engine **out) {
    if (!g || !p || !out) return LT_E_ARG;
    lt_config c;
    if (cfg) c = *cfg;
    else lt_config_default(&c);
    std::unique_ptr<lt_engine> e(new lt_engine);
    lt_status s = e->init(g, p, &c);
    if (s != LT_OK) return s;
    *out = e.release();
    return LT_OK;
}

void lt_engine_destroy(lt_engine *e) { delete e; }

// Walker ids are 64-bit: w * runs walkers must not wrap.
static bool walkers_fit(uint64_t w, uint32_t runs) {
    if (w > UINT64_MAX / runs) {
        set_error("w * runs = %llu * %u walkers exceed 64-bit walker ids", (unsigned long long)w, runs);
        return false;
    }
    return true;
}

// The termination probability of a PPR walk (app/ppr.cu:22-42), in (0, 1].
static bool prob_ok(float prob) {
    if (!(prob > 0.0f && prob <= 1.0f)) {
        set_error("prob must be in (0, 1] (got %g)", (double)prob);
        return false;
    }
    return true;
}

lt_status lt_run_pagerank(lt_engine *e, uint64_t w, uint32_t l, uint32_t runs, uint32_t *counts_out,
                          uint32_t *endpoints_out) {
    if (!e || w == 0 || runs == 0 || !walkers_fit(w, runs)) return LT_E_ARG;
    RunSpec rs;
    rs.app = APP_PAGERANK;
    rs.total = w * runs;
    rs.length = l;
    rs.nominal = w * (uint64_t)l * runs;
    return e->run(rs, counts_out, endpoints_out, nullptr);
}

lt_status lt_run_ppr(lt_engine *e, uint64_t w, uint32_t source, float prob, uint32_t *counts_out,
                     uint32_t *endpoints_out) {
    if (!e || w == 0 || !prob_ok(prob)) return LT_E_ARG;
    if (source >= e->V) {
        set_error("source %u >= number of vertices %u", source, e->V);
        return LT_E_ARG;
    }
    RunSpec rs;
    rs.app = APP_PPR;
    rs.total = w;
    rs.source = source;
    rs.prob = prob;
    rs.nominal = (uint64_t)((double)w / (double)prob);
    return e->run(rs, counts_out, endpoints_out, nullptr);
}

lt_status lt_run_genericwalk(lt_engine *e, uint64_t w, uint32_t l, uint32_t runs, uint32_t *endpoints_out,
                             uint32_t *moves_out) {
    if (!e || w == 0 || runs == 0 || !walkers_fit(w, runs)) return LT_E_ARG;
    RunSpec rs;
    rs.app = APP_GENERIC;
    rs.total = w * runs;
    rs.length = l;
    rs.nominal = w * (uint64_t)l * runs;
    return e->run(rs, nullptr, endpoints_out, moves_out);
}

lt_status lt_run_genericwalk_paths(lt_engine *e, uint64_t w, uint32_t l, uint32_t runs, uint32_t *paths_out) {
    if (!e || w == 0 || runs == 0 || !paths_out || !walkers_fit(w, runs)) return LT_E_ARG;
    RunSpec rs;
    rs.app = APP_GENERIC;
    rs.total = w * runs;
    rs.length = l;
    rs.nominal = w * (uint64_t)l * runs;
    rs.paths_out = paths_out;
    return e->run(rs, nullptr, nullptr, nullptr);
}

lt_status lt_run_ppr_batch(lt_engine *e, uint32_t nq, const uint32_t *sources, const uint64_t *counts, float prob,
                           uint32_t *endpoints_out) {
    if (!e || nq == 0 || !sources || !counts || !prob_ok(prob)) return LT_E_ARG;
    uint64_t total = 0;
    for (uint32_t i = 0; i < nq; i++) {
        if (sources[i] >= e->V) {
            set_error("query %u source %u >= number of vertices %u", i, sources[i], e->V);
            return LT_E_ARG;
        }
        if (counts[i] > UINT64_MAX - total) {
            set_error("the queries' walk counts exceed 64-bit walker ids");
            return LT_E_ARG;
        }
        total += counts[i];
    }
    if (total == 0) {
        set_error("the queries hold no walks");
        return LT_E_ARG;
    }
    RunSpec rs;
    rs.app = APP_PPR;
    rs.total = total;
    rs.prob = prob;
    rs.q_src = sources;
    rs.q_cnt = counts;
    rs.nq = nq;
    rs.nominal = (uint64_t)((double)total / (double)prob);
    return e->run(rs, nullptr, endpoints_out, nullptr);
}

lt_status lt_device_h2d_bandwidth(uint64_t bytes, uint32_t reps, double *gbps_out) {
    if (bytes == 0 || reps == 0) return LT_E_ARG;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        (void)hipGetLastError();
        set_error("no HIP device visible");
        return LT_E_NODEVICE;
    }
    uint8_t *h = nullptr, *d = nullptr;
    LT_HIP(hipHostMalloc((void **)&h, bytes, hipHostMallocPortable));
    std::unique_ptr<uint8_t, void (*)(uint8_t *)> gh(h, [](uint8_t *x) { (void)hipHostFree(x); });
    memset(h, 1, bytes);
    LT_TRY(dmalloc(&d, bytes));
    std::unique_ptr<uint8_t, void (*)(uint8_t *)> gd(d, [](uint8_t *x) { (void)hipFree(x); });
    hipStream_t s;
    LT_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    std::unique_ptr<ihipStream_t, void (*)(hipStream_t)> gs(s, [](hipStream_t x) { (void)hipStreamDestroy(x); });
    LT_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s));  // warm
    LT_HIP(hipStreamSynchronize(s));
    hipEvent_t e0, e1;
    LT_HIP(hipEventCreate(&e0));
    LT_HIP(hipEventCreate(&e1));
    LT_HIP(hipEventRecord(e0, s));
    for (uint32_t r = 0; r < reps; r++) LT_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s));
    LT_HIP(hipEventRecord(e1, s));
    LT_HIP(hipEventSynchronize(e1));
    float ms = 0;
    LT_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (gbps_out) *gbps_out = (double)bytes * reps / (ms * 1e-3) / 1e9;
    return LT_OK;
}

lt_status lt_device_copy_bandwidth(uint64_t bytes, uint32_t reps, double *gbps_out) {
    if (bytes < 16 || reps == 0) return LT_E_ARG;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        (void)hipGetLastError();
        set_error("no HIP device visible");
        return LT_E_NODEVICE;
    }
    const uint64_t n16 = bytes / 16;
    uint4 *a = nullptr, *b = nullptr;
    LT_TRY(dmalloc(&a, n16));
    std::unique_ptr<uint4, void (*)(uint4 *)> ga(a, [](uint4 *x) { (void)hipFree(x); });
    LT_TRY(dmalloc(&b, n16));
    std::unique_ptr<uint4, void (*)(uint4 *)> gb(b, [](uint4 *x) { (void)hipFree(x); });
    fill_u32_kernel<<<1024, 256>>>((uint32_t *)a, n16 * 4, 1u);
    copy_probe_kernel<<<4096, 256>>>(a, b, n16);  // warm
    LT_HIP(hipGetLastError());
    LT_HIP(hipDeviceSynchronize());
    hipEvent_t e0, e1;
    LT_HIP(hipEventCreate(&e0));
    LT_HIP(hipEventCreate(&e1));
    LT_HIP(hipEventRecord(e0, nullptr));
    for (uint32_t r = 0; r < reps; r++) copy_probe_kernel<<<4096, 256>>>(a, b, n16);
    LT_HIP(hipEventRecord(e1, nullptr));
    LT_HIP(hipEventSynchronize(e1));
    float ms = 0;
    LT_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (gbps_out) *gbps_out = 2.0 * 16.0 * (double)n16 * reps / (ms * 1e-3) / 1e9;
    return LT_OK;
}

lt_status lt_engine_stats(const lt_engine *e, lt_stats *out) {
    if (!e || !out) return LT_E_ARG;
    *out = e->st;
    return LT_OK;
}

lt_status lt_engine_fetch_partition(lt_engine *e, uint32_t part, uint64_t *rowptr_out, uint32_t *col_out) {
    if (!e || !rowptr_out || !col_out || part >= e->P) {
        lt::set_error("lt_engine_fetch_partition: null argument or partition %u out of range", part);
        return LT_E_ARG;
    }
    return e->fetch_partition(part, rowptr_out, col_out);
}

uint64_t lt_engine_trace_len(const lt_engine *e) { return e ? e->trace.size() : 0; }

lt_status lt_engine_trace_state(const lt_engine *e, uint32_t *counts_out, uint8_t *resident_out, uint32_t *age_out,
                                int32_t *recent_out, uint64_t cap) {
    if (!e) return LT_E_ARG;
    const uint64_t n = std::min<uint64_t>(cap, e->trace.size());
    const size_t P = e->P;
    if (counts_out && n) memcpy(counts_out, e->trace_counts.data(), n * P * sizeof(uint32_t));
    if (resident_out && n) memcpy(resident_out, e->trace_resident.data(), n * P);
    if (age_out && n) memcpy(age_out, e->trace_age.data(), n * P * sizeof(uint32_t));
    if (recent_out && n) memcpy(recent_out, e->trace_recent.data(), n * kRecent * sizeof(int32_t));
    return LT_OK;
}

// Training samples from a traced run.  Decisions are the records of kind 0,
// 1 and 3 (kind 2 = pipelined work inside a load).  Output 0's label is the
// next decision's partition when that decision needed host data (kind 1 or
// 3).  Output 1's label, with use_keep = 1, is each partition's share of the
// walkers processed by all later records, times P; with use_keep = 2 it is
// Belady's quantity, how soon the partition is next needed in HBM: with d
// the number of later records of kind 0, 2 or 3 up to and including its next
// one, 1 - ln(1 + d) / ln(1 + n) (0 when it is never needed again), so that
// evicting the lowest score evicts the partition needed farthest ahead.
lt_status lt_scorer_add_trace(lt_scorer *s_, const lt_engine *e, int32_t use_next, int32_t use_keep,
                              uint64_t *added) {
    if (!s_ || !e) return LT_E_ARG;
    lt::Scorer *s = (lt::Scorer *)s_;
    const uint32_t P = e->P;
    if (scorer_partitions(s) != P) {
        set_error("scorer built for %u partitions, engine has %u", scorer_partitions(s), P);
        return LT_E_ARG;
    }
    const auto &tr = e->trace;
    const size_t n = tr.size();
    if (n == 0) {
        set_error("the engine holds no trace (set record_trace)");
        return LT_E_ARG;
    }
    if (use_keep < 0 || use_keep > 2) {
        set_error("use_keep must be 0, 1 or 2 (got %d)", use_keep);
        return LT_E_ARG;
    }
    // suffix traffic per partition, and the position of each partition's next
    // HBM demand, built from the back
    std::vector<double> future(P, 0.0);
    std::vector<uint64_t> next_demand(P, UINT64_MAX);
    std::vector<std::vector<float>> keep_at(n);
    double total = 0;
    uint64_t demands_after = 0;  // records of kind 0, 2, 3 after position j
    int32_t next_need = -1;
    std::vector<int32_t> next_of(n, -1);
    const double ln_n = log1p((double)n);
    for (size_t j = n; j-- > 0;) {
        const auto &r = tr[j];
        if (r.kind != 2) {
            next_of[j] = next_need;
            keep_at[j].resize(P);
            for (uint32_t q = 0; q < P; q++) {
                if (use_keep == 2) {
                    const double d = next_demand[q] == UINT64_MAX ? -1.0 : (double)(demands_after - next_demand[q] + 1);
                    keep_at[j][q] = d < 0 ? 0.f : (float)(1.0 - log1p(d) / ln_n);
                } else {
                    keep_at[j][q] = total > 0 ? (float)(future[q] / total * P) : 0.f;
                }
            }
            next_need = (r.kind == 1 || r.kind == 3) ? r.partition : -1;
        }
        if (r.kind != 1) {  // needed in HBM at this record; numbered from the back
            demands_after++;
            next_demand[r.partition] = demands_after;
        }
        future[r.partition] += (double)r.walkers;
        total += (double)r.walkers;
    }
    Samples &d = scorer_samples(s);
    uint64_t cnt = 0;
    for (size_t j = 0; j < n; j++) {
        const auto &r = tr[j];
        if (r.kind == 2) continue;
        const bool has_next = use_next && next_of[j] >= 0;
        const bool has_keep = use_keep != 0;
        if (!has_next && !has_keep) continue;
        d.counts.insert(d.counts.end(), e->trace_counts.begin() + j * P, e->trace_counts.begin() + (j + 1) * P);
        d.resident.insert(d.resident.end(), e->trace_resident.begin() + j * P, e->trace_resident.begin() + (j + 1) * P);
        d.age.insert(d.age.end(), e->trace_age.begin() + j * P, e->trace_age.begin() + (j + 1) * P);
        d.recent.insert(d.recent.end(), e->trace_recent.begin() + j * kRecent,
                        e->trace_recent.begin() + (j + 1) * kRecent);
        d.cur.push_back(r.partition);
        d.next.push_back(has_next ? next_of[j] : -1);
        d.w_total.push_back((float)e->w_run);
        d.keep.insert(d.keep.end(), keep_at[j].begin(), keep_at[j].end());
        d.wa.push_back(has_next ? 1.f : 0.f);
        d.wb.push_back(has_keep ? 1.f : 0.f);
        d.n++;
        cnt++;
    }
    if (added) *added = cnt;
    return LT_OK;
}

lt_status lt_engine_set_scorer(lt_engine *e, lt_scorer *s) {
    if (!e) return LT_E_ARG;
    if (s && scorer_partitions((lt::Scorer *)s) != e->P) {  // its state and scores are sized by P
        set_error("scorer built for %u partitions, engine has %u", scorer_partitions((lt::Scorer *)s), e->P);
        return LT_E_ARG;
    }
    LT_TRY(e->infer_collect());  // an inference of the previous scorer still in flight
    e->scores_valid = false;
    e->scorer = (lt::Scorer *)s;
    return LT_OK;
}

lt_status lt_engine_trace(const lt_engine *e, lt_trace_rec *out, uint64_t cap) {
    if (!e || !out) return LT_E_ARG;
    const uint64_t n = std::min<uint64_t>(cap, e->trace.size());
    if (n) memcpy(out, e->trace.data(), n * sizeof(lt_trace_rec));
    return LT_OK;
}

}  // extern "C"
