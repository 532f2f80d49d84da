/*
 * lighttraffic.h — C ABI of the MI355X random-walk engine (liblighttraffic.so).
 *
 * The reference exposes its engine as compile-time C++ templates
 * (LightTraffic<AppManager, App, Walker>, src/LightTraffic.h:11-45) driven by
 * three CUDA programs (app/pagerank.cu, app/ppr.cu, app/genericwalk.cu) and a
 * flag parser (src/Config.h:38-89).  This header replaces that plug-in surface
 * with opaque handles, plain pointers and sizes: the three walk kinds are a
 * closed set selected by the entry point, and every result is written into a
 * caller-owned host buffer.  No torch or HIP types appear in any signature.
 *
 * Every entry point returns an lt_status; the library never calls exit().
 * Handles are not thread-safe; two engines may coexist in one process.
 */
#ifndef LIGHTTRAFFIC_H
#define LIGHTTRAFFIC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  Replace the reference's assert()/exit()/throw paths:
 * gpuCall -> exit(code) (src/utils/GPUCall.h:3-11), assert(fp) on a missing
 * .part file (src/partition/PartitionStrategy.h:80), runtime_error for a vertex
 * larger than a partition (src/partition/PartitionStrategy.h:201-205). */
typedef enum {
    LT_OK = 0,
    LT_E_IO = 1,               /* file missing, short read or write            */
    LT_E_FORMAT = 2,           /* bcsr or .part content is inconsistent        */
    LT_E_OVERSIZE_VERTEX = 3,  /* one vertex does not fit the partition size   */
    LT_E_NOMEM = 4,            /* device or pinned host allocation failed      */
    LT_E_HIP = 5,              /* HIP runtime error (message in lt_last_error) */
    LT_E_ARG = 6,              /* invalid argument                             */
    LT_E_NODEVICE = 7,         /* no HIP device visible                        */
    LT_E_INTERNAL = 8          /* engine invariant violated (bug)              */
} lt_status;

typedef struct lt_graph lt_graph;         /* bcsr CSR in pinned host memory  */
typedef struct lt_partition lt_partition; /* contiguous vertex-range split   */
typedef struct lt_engine lt_engine;       /* GPU engine bound to one graph   */
typedef struct lt_scorer lt_scorer;       /* learned partition scorer (MLP)  */

/* Partition scheduler policies.  GREEDY and FIFO restate the reference's two
 * compile-time variants of Scheduler::nextPartition (src/Scheduler.h:368-388
 * and :393-416, selected by -DFIFO_SCHE); LEARNED is new: an MLP scores the
 * partitions and the engine falls back to GREEDY when it has no valid score. */
typedef enum { LT_SCHED_GREEDY = 0, LT_SCHED_FIFO = 1, LT_SCHED_LEARNED = 2 } lt_sched_policy;

/* How far one residency of a walker reaches.  PARTITION walks while the
 * current vertex lies in the partition being processed, as the reference's
 * randomWalkKernel loop does (src/WalkManager.h:239-240).  RESIDENT keeps
 * walking while the vertex lies in any partition held in the HBM cache. */
typedef enum { LT_SPAN_PARTITION = 0, LT_SPAN_RESIDENT = 1 } lt_span;

typedef struct {
    uint32_t num_device_partitions; /* -p: HBM cache slots (0 = all partitions)   */
    uint64_t num_device_walkers;    /* -b: walkers resident at once (0 = all)     */
    uint64_t zerocopy_threshold;    /* walkers below which a non-resident
                                       partition is read over PCIe; UINT64_MAX =
                                       default: partition_bytes/256 (reference)
                                       for GREEDY/FIFO, partition_bytes/128 (byte
                                       break-even of 64 B PCIe reads) for LEARNED */
    uint64_t seed;                  /* Philox key                                 */
    int32_t  policy;                /* lt_sched_policy                            */
    int32_t  pipeline;              /* 1: compute other partitions while loading  */
    int32_t  span;                  /* lt_span                                    */
    int32_t  event_timers;          /* 1: per-phase HIP event timers              */
    int32_t  record_trace;          /* 1: keep the scheduler decision trace       */
    int32_t  prefetch_depth;        /* LEARNED: partitions copied ahead (0 = off)  */
    int32_t  prefetch_stick_pct;    /* LEARNED: walk a prefetched partition first
                                       when its queue is >= this % of the largest */
    int32_t  device_schedule;       /* 1: GREEDY decisions, partition copies and
                                       walks all run on the GPU, 32 iterations
                                       per hipGraph launch (no host round trip) */
    int32_t  slice_chunks;          /* chunks of 4096 walkers per walk launch
                                       (0 = 2048); device_schedule: unused      */
    int32_t  zc_batch;              /* span RESIDENT: walk up to this many
                                       partitions below the zero-copy threshold
                                       over PCIe in one launch (at most 16;
                                       0 = 16 for LEARNED, 1 otherwise)         */
    int32_t  rowptr_resident;       /* 1: keep the whole rowptr array (8 B per
                                       vertex) in HBM, taken from the partition
                                       cache's slots, so a zero-copied move reads
                                       only its column entry over PCIe (default 0:
                                       off)                                     */
    int32_t  pack_partitions;       /* 1: copy partition columns gap-coded
                                       (lt_partition_pack) and unpack them in
                                       HBM; 2: copy them as 16-edge
                                       frame-of-reference blocks that stay
                                       packed in the cache slot and are read in
                                       place, so slots are smaller and the
                                       cache holds more partitions.  A
                                       partition that does not get smaller is
                                       copied as is (default 0: off;
                                       device_schedule: 2 only, 1 is ignored)   */
    int32_t  prefetch_plain;        /* LEARNED: prefetch targets and victims.
                                       0: the scorer's "next" output picks the
                                       targets, its "keep" output the victims;
                                       1: no scorer: the largest queues that
                                       need host data are the targets and the
                                       plain rule picks victims (an ablation
                                       that shows what the network adds); 2:
                                       largest-queue targets, the scorer's
                                       victims (default 0)                     */
    int32_t  evict_shortlist;       /* LEARNED victims: the scorer chooses among
                                       this many least recently walked of the
                                       resident partitions with the fewest
                                       queued walkers (0 = among all of them;
                                       1 = least recently walked, no scorer
                                       choice; default 0)                      */
    int32_t  walkers_per_thread;    /* walkers one walk-kernel lane holds: 0 =
                                       by launch size (4, 2 or 1, host
                                       scheduling; 1 in the device loop), or
                                       exactly 1, 2 or 4                        */
    int32_t  resident_batch;        /* span RESIDENT: a launch over a resident
                                       partition's queue also takes the largest
                                       other resident queues of at least 4096
                                       walkers, up to this many partitions (at
                                       most 16; 0 = 16 for LEARNED, 1
                                       otherwise)                              */
    int32_t  learned_next;          /* LEARNED, with a scorer: when the greedy
                                       choice (Scheduler::nextPartition,
                                       src/Scheduler.h:368-388) needs host
                                       data, walk instead the partition with
                                       the highest "next" score among those
                                       whose queue is at least this % of the
                                       greedy choice's (0 = off, the default;
                                       1-100)                                  */
    int32_t  append_mode;           /* how the walk appends leaving walkers to
                                       their queues: 0 (default) one queue
                                       reservation per (workgroup, queue),
                                       walkers staged in LDS for coalesced
                                       stores; 1 the same reservations, each
                                       walker stored from its own lane (the
                                       reference's NO_RANDOM_ACCESS_REDUCE
                                       build, src/WalkManager.h:98-118); 2 one
                                       reservation per walker (its
                                       ONE_LEVEL_QUEUE build,
                                       src/WalkManager.h:195-209).  Figure 12's
                                       ablations; results do not change      */
} lt_config;

/* Statistics of the last run call.  The first eight fields mirror the lines
 * that Scheduler::timing prints (src/Scheduler.h:440-453). */
typedef struct {
    double   running_ms;            /* "Running Time" (walk phase only)          */
    uint64_t iterations;            /* scheduler iterations                      */
    uint64_t explicit_iterations;   /* iterations that used an HBM-resident copy */
    double   graph_loading_ms;  uint64_t graph_loading_calls;
    double   computing_ms;      uint64_t computing_calls;
    double   zerocopy_ms;       uint64_t zerocopy_calls;
    double   walker_loading_ms; uint64_t walker_loading_calls;
    double   walker_evicting_ms; uint64_t walker_evicting_calls;
    uint64_t h2d_bytes_explicit;    /* partition bytes copied host->HBM          */
    uint64_t h2d_bytes_zerocopy;    /* estimated bytes read over PCIe in place   */
    uint64_t moves;                 /* neighbour moves (rowptr pair + col read)  */
    uint64_t teleports;             /* pagerank jumps to a uniform vertex        */
    uint64_t visits;                /* visit-counter increments                  */
    uint64_t residencies;           /* walker records read from a queue          */
    uint64_t zerocopy_moves;        /* moves whose graph reads crossed PCIe      */
    uint64_t nominal_steps;         /* w*l*runs (pagerank, genericwalk), w/prob  */
    uint64_t kernel_launches;       /* walk-kernel launches                      */
    double   walk_kernel_ms;        /* summed walk-kernel time (event timers)    */
    uint64_t learned_decisions;     /* evictions chosen by the MLP keep score    */
    uint64_t fallback_decisions;    /* evictions that fell back to the plain rule */
    uint64_t prefetch_issued;       /* partitions copied ahead of their turn     */
    uint64_t prefetch_hits;         /* prefetched partitions picked next         */
    uint64_t h2d_bytes_unpacked;    /* the bytes the explicit copies stand for
                                       (= h2d_bytes_explicit without packing)   */
    uint64_t packed_image_bytes;    /* packed column images copied (part of
                                       h2d_bytes_explicit)                      */
    uint64_t packed_col_bytes;      /* column bytes those images unpacked to    */
    uint64_t cache_slots;           /* partitions the HBM cache holds at once   */
    uint64_t zerocopy_launches;     /* walk launches whose source queues were
                                       read over PCIe (part of kernel_launches) */
    double   walk_kernel_zc_ms;     /* their summed time (part of walk_kernel_ms) */
    uint64_t resident_batched;      /* resident queues walked in another
                                       resident queue's launch (resident_batch) */
    uint64_t next_overrides;        /* decisions where learned_next walked
                                       another partition than the greedy one   */
} lt_stats;

/* ---- errors ------------------------------------------------------------- */
const char *lt_status_str(lt_status s);
const char *lt_last_error(void);            /* detail of the last failure      */
void lt_config_default(lt_config *cfg);

/* ---- graph (src/GraphLoader.h:30-48 DiskGraph, :127-158 CPUGraph) -------- */
/* bcsr: u32 num_node, u64 num_edge, u64 rowptr[V+1], u32 col[E]. */
lt_status lt_graph_open(const char *path, lt_graph **out);
lt_status lt_graph_from_csr(uint32_t num_node, uint64_t num_edge, const uint64_t *rowptr,
                            const uint32_t *col, lt_graph **out);
/* R-MAT generator on the GPU: ef*2^scale edge samples with quadrant
 * probabilities a,b,c (d = 1-a-b-c) drawn from Philox(seed), then self-loops
 * dropped, isolated vertices removed, edges symmetrised, sorted and
 * deduplicated — the pipeline of tools/convert_from_txt.cc:66-218. */
lt_status lt_graph_generate_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                                 uint64_t seed, lt_graph **out);
lt_status lt_graph_save(const lt_graph *g, const char *path);
uint32_t lt_graph_num_node(const lt_graph *g);
uint64_t lt_graph_num_edge(const lt_graph *g);
const uint64_t *lt_graph_rowptr(const lt_graph *g);
const uint32_t *lt_graph_col(const lt_graph *g);
void lt_graph_free(lt_graph *g);

/* ---- partitions (src/partition/PartitionStrategy.h:65-338) -------------- */
/* Greedy contiguous split, (nodes+1)*8 + edges*4 <= max_bytes per partition
 * (EqualSizeStrategy::init, PartitionStrategy.h:189-235). */
lt_status lt_partition_build(const lt_graph *g, uint64_t max_bytes, lt_partition **out);
/* .part: u16 P, u64 partition_bytes, u32 node_offset[P+1], u64 edge_offset[P+1]
 * written to <graph_path>.part (materialize, PartitionStrategy.h:325-337) and
 * read back (MaterializedStrategy, PartitionStrategy.h:76-97). */
lt_status lt_partition_save(const lt_partition *p, const char *graph_path);
lt_status lt_partition_load(const char *graph_path, lt_partition **out);
lt_status lt_partition_from_offsets(uint32_t num_partition, uint64_t partition_bytes,
                                    const uint32_t *node_offset, const uint64_t *edge_offset,
                                    lt_partition **out);
uint32_t lt_partition_count(const lt_partition *p);
uint64_t lt_partition_bytes(const lt_partition *p);
const uint32_t *lt_partition_node_offsets(const lt_partition *p);
const uint64_t *lt_partition_edge_offsets(const lt_partition *p);
/* Host findPartition (PartitionStrategy.h:151-168): partition holding v. */
uint32_t lt_partition_find(const lt_partition *p, uint32_t v);
void lt_partition_free(lt_partition *p);
/* Packed image of partition part's column slice, the form in which an engine
 * with pack_partitions = 1 copies it to HBM (replaces the raw col copy of
 * GPUGraph::loadAsync, src/GraphLoader.h:193-215; the rowptr slice is copied
 * as is).  For m edges in nch = ceil(m/64) chunks of 64 (partition-local edges
 * 64c .. 64c+63), each of four 16-edge blocks, with vbits = bits of V-1 (at
 * least 1):
 *   u32 meta[2*nch]: meta[2c] = first word of chunk c in words[]; byte j of
 *       meta[2c+1] = code of block j of chunk c: 0..32 = bit width of the
 *       block's gaps, 255 = block stored absolute;
 *   u32 words[]: each chunk's bit stream (LSB first, from its first word),
 *       then two zero words.
 * An edge is an anchor when it is its row's first, its chunk's first, or its
 * block's code is 255; an anchor holds its column in vbits bits, any other
 * edge holds the gap to the previous column in its block's width.  *bytes_out
 * gets the image size; the image is written when out is not NULL and cap is
 * large enough. */
lt_status lt_partition_pack(const lt_graph *g, const lt_partition *p, uint32_t part, void *out, uint64_t cap,
                            uint64_t *bytes_out);
/* Block image of partition part's column slice, the form in which an engine
 * with pack_partitions = 2 copies it to HBM and keeps it there, read in
 * place by the walk (replaces the raw col slice of the cache slot,
 * GPUGraph::loadAsync, src/GraphLoader.h:193-215).  For m edges in
 * nblk = ceil(m/16) blocks of 16 (partition-local edges 16k .. 16k+15; the
 * last block padded with its base):
 *   u32 hdr[2*(nblk+1)]: hdr[2k] = base of block k (its smallest column),
 *       hdr[2k+1] = its first bit in words[]; hdr[2*nblk] = 0,
 *       hdr[2*nblk+1] = total bits;
 *   u32 words[]: the blocks' bit streams back to back (LSB first), then two
 *       zero words.
 * Block k stores column - base of each of its 16 edges in
 * w = (hdr[2k+3] - hdr[2k+1]) / 16 bits.  Same size query and write rules as
 * lt_partition_pack.  No counterpart in the reference. */
lt_status lt_partition_pack_blocks(const lt_graph *g, const lt_partition *p, uint32_t part, void *out,
                                   uint64_t cap, uint64_t *bytes_out);
/* HBM cache layout an engine without device_schedule gets for partition table
 * p: *slot_bytes per slot and *slots slots for num_device_partitions raw slots
 * (0: one per partition), before rowptr_resident takes its share.  With
 * pack_partitions 0 or 1 a slot is S rounded up to 256 bytes (the reference's
 * slot, src/GraphLoader.h:175-176); with 2 it is the largest rowptr slice +
 * block image over the partitions (256-aligned) and the cache keeps the same
 * bytes, so it holds more slots (at most one per partition).  Lets a caller
 * size the cache before creating an engine; engines report the same count in
 * lt_stats.cache_slots.  No counterpart in the reference. */
lt_status lt_partition_slot_layout(const lt_graph *g, const lt_partition *p, int32_t pack_partitions,
                                   uint32_t num_device_partitions, uint64_t *slot_bytes, uint32_t *slots);
/* Greedy contiguous split sized for block slots: each partition's rowptr
 * slice ((nodes+1)*8 bytes) plus its block image (lt_partition_pack_blocks;
 * its raw columns when the image is not smaller) fits slot_bytes, and the
 * next vertex would not.  The reference splits by raw bytes only
 * (EqualSizeStrategy::init, PartitionStrategy.h:189-235), so its block
 * images vary in size and a block slot, sized by the largest, goes partly
 * unused; balanced images let the same HBM cache hold more of the graph.
 * partition_bytes is set to the largest partition's raw size
 * ((nodes+1)*8 + edges*4), so raw slots hold these partitions as well.
 * LT_E_OVERSIZE_VERTEX if one vertex does not fit.  No counterpart in the
 * reference. */
lt_status lt_partition_build_blocks(const lt_graph *g, uint64_t slot_bytes, lt_partition **out);
/* An engine with pack_partitions = 1 or 2 builds the packed images of its
 * partition table once (pinned host memory; on R-MAT about half the graph's
 * column bytes for 1, 0.6 for 2) and caches them on the graph for later
 * engines.  This drops the
 * cache; engines already using an image keep it until they are destroyed.
 * No counterpart in the reference. */
void lt_graph_drop_packed(lt_graph *g);

/* ---- engine (src/LightTraffic.h:19-44, src/Device.h:22-59) -------------- */
lt_status lt_engine_create(const lt_graph *g, const lt_partition *p, const lt_config *cfg,
                           lt_engine **out);
void lt_engine_destroy(lt_engine *e);

/* pagerank (app/pagerank.cu:26-44): w walkers from uniform starts, l steps,
 * teleport with probability 0.15, one visit counted per step.  counts_out
 * (u32[V], may be NULL) receives the summed visit counts of all runs;
 * endpoints_out (u32[w*runs], may be NULL) each walker's last vertex. */
lt_status lt_run_pagerank(lt_engine *e, uint64_t w, uint32_t l, uint32_t runs, uint32_t *counts_out,
                          uint32_t *endpoints_out);
/* ppr (app/ppr.cu:22-42, src/walk/SingleSource.h:6-20): w walkers from
 * source, length 2+Geom(prob) drawn up front, endpoint counted. endpoints_out
 * (u32[w], may be NULL) receives each walker's final vertex. */
lt_status lt_run_ppr(lt_engine *e, uint64_t w, uint32_t source, float prob, uint32_t *counts_out,
                     uint32_t *endpoints_out);
/* genericwalk (app/genericwalk.cu:22-32): fixed-length uniform walks that stop
 * early at a vertex without neighbours.  The reference records nothing; this
 * entry point can return each walker's final vertex (u32[w*runs]) and the
 * number of moves it made (u32[w*runs]). */
lt_status lt_run_genericwalk(lt_engine *e, uint64_t w, uint32_t l, uint32_t runs,
                             uint32_t *endpoints_out, uint32_t *moves_out);
/* genericwalk with the whole path of every walker (no counterpart in the
 * reference's programs; its layout is that of the reference's unused path
 * samplers, src/walk/NodeSampler.h:36-49 and src/walk/PathSampler.h:6-19):
 * paths_out is u32[w*runs][l+1], row = walker id, column k = the vertex after
 * k moves (column 0 the start), 0xFFFFFFFF after a walk that stopped early at
 * a vertex without neighbours. */
lt_status lt_run_genericwalk_paths(lt_engine *e, uint64_t w, uint32_t l, uint32_t runs, uint32_t *paths_out);
/* Batched PPR (no counterpart in the reference): nq queries, query q starts
 * counts[q] walkers at sources[q]; walker ids are numbered query by query.
 * endpoints_out: u32[sum counts] in walker-id order (may be NULL). */
lt_status lt_run_ppr_batch(lt_engine *e, uint32_t nq, const uint32_t *sources, const uint64_t *counts,
                           float prob, uint32_t *endpoints_out);
lt_status lt_engine_stats(const lt_engine *e, lt_stats *out);
/* Partition part as the engine's partition copy writes it into an HBM cache
 * slot (packed or not, per the engine's config), copied back to the host:
 * rowptr_out u64[n+1], col_out u32[m].  Leaves the cache untouched (no
 * counterpart in the reference; it lets a caller check a slot's bytes). */
lt_status lt_engine_fetch_partition(lt_engine *e, uint32_t part, uint64_t *rowptr_out, uint32_t *col_out);
/* Batched-PPR query table (no counterpart in the reference): query q draws
 * Philox(seed; q, 0, TAG_QUERY) = (x0,x1,x2,x3); its source is the row that
 * owns edge bounded(x0,x1,E) (sources proportional to out-degree), and its
 * walk count is floor(lo * (hi/lo)^(x2/2^32)) (log-uniform in [lo, hi)). */
lt_status lt_make_ppr_queries(const lt_graph *g, uint32_t nq, uint64_t lo, uint64_t hi, uint64_t seed,
                              uint32_t *sources_out, uint64_t *counts_out);

/* Autotuner of batched PPR (no counterpart in the reference, whose knobs are
 * compile-time, src/Type.h:5-22, and which runs one source per program,
 * src/walk/SingleSource.h:6-20).  Times engine configurations on every
 * sample_every-th query of the batch (a warm second job each): partition
 * size sizes_mb[0..nsizes) (NULL: 16, 32 and 64 MB) with cache_bytes of
 * cache, span partition or resident,
 * host or device scheduling, 2048 or 8192 chunks per walk launch, zero-copy
 * and resident batches of 1 or 16 (zc_batch = resident_batch; host
 * scheduling, span resident), raw / gap-coded /
 * block copies (gap-coded needs host scheduling), the block copies also on
 * the split sized for block slots of S (lt_partition_build_blocks).  Then it
 * times the three fastest and the fastest of each copy mode (block copies on
 * either split) again on a sample four times
 * larger, then the fastest of those with each fixed number of walkers per
 * walk-kernel lane (1, 2, 4) on that sample, and returns the fastest of
 * stages 2 and 3 in *best_out.  Every trial is written to trials_out (up to
 * trials_cap; *ntrials_out = their number).  The queue chunk (4096 walkers)
 * is fixed at compile time; slice_chunks is the launch size searched.
 * Results of the batch do not depend on the configuration. */
typedef struct {
    uint32_t stage;            /* 1: timed on the sample, 2: again on the larger
                                  one, 3: stage 2's pick with a fixed WPT        */
    uint32_t partition_mb;     /* partition size S in MB                           */
    uint32_t slots;            /* num_device_partitions = cache_bytes / S (<= P)   */
    uint32_t partitions;       /* P at this size                                   */
    uint32_t cache_slots;      /* partitions the cache held (lt_stats.cache_slots) */
    int32_t  span;             /* the lt_config fields of the configuration        */
    int32_t  device_schedule;
    int32_t  slice_chunks;
    int32_t  zc_batch;
    int32_t  pack_partitions;
    int32_t  walkers_per_thread;
    int32_t  split;            /* 0: lt_partition_build(S); 1 (block copies only):
                                  lt_partition_build_blocks(S), slots =
                                  cache_bytes / its partition_bytes             */
    double   running_ms;       /* the warm job's scheduler-loop time               */
    uint64_t h2d_bytes;        /* its explicit + zero-copied H2D bytes             */
    uint64_t iterations;       /* its scheduler iterations                         */
} lt_tune_trial;
lt_status lt_tune_ppr_batch(const lt_graph *g, uint32_t nq, const uint32_t *sources, const uint64_t *counts,
                            float prob, uint64_t cache_bytes, const uint32_t *sizes_mb, uint32_t nsizes,
                            uint32_t sample_every, lt_tune_trial *best_out, lt_tune_trial *trials_out,
                            uint32_t trials_cap, uint32_t *ntrials_out);

/* Scheduler decision trace of the last run (record_trace=1): one record per
 * decision.  kind: 0 partition already resident, 1 zero-copy, 2 pipelined
 * compute during a load, 3 partition loaded (evicted = the slot's previous
 * partition or -1).  walkers = queue length of the chosen partition. */
typedef struct {
    int32_t partition;
    int32_t kind;
    int32_t loaded;
    int32_t evicted;
    uint64_t walkers;
} lt_trace_rec;
uint64_t lt_engine_trace_len(const lt_engine *e);
lt_status lt_engine_trace(const lt_engine *e, lt_trace_rec *out, uint64_t cap);
/* State seen by each traced decision: queue lengths (u32[len*P]), residency
 * (u8[len*P]), partition ages (u32[len*P]: walks of resident partitions since
 * the partition's own last one, counted from the engine's creation) and the
 * last 8 loaded partitions (i32[len*8], most recent first, -1 = none); any
 * output may be NULL.  cap counts records. */
lt_status lt_engine_trace_state(const lt_engine *e, uint32_t *counts_out, uint8_t *resident_out,
                                uint32_t *age_out, int32_t *recent_out, uint64_t cap);

/* ---- learned scheduler (no counterpart in the reference) ---------------- */
/* Per-partition MLP, 16 features -> 256 -> 256 -> 2, BF16 weights, FP32
 * accumulation, trained and evaluated by HIP kernels.  Features: queue length
 * (three scalings), residency, chosen now, loaded 0..6 decisions ago, age
 * a/(a+P), vertex and edge shares, queued walkers / W.  Output 0 scores the
 * partition the scheduler will need from the host next (used to prefetch it
 * while the current partition is walked); output 1 predicts how soon the
 * partition is needed in HBM again (use_keep = 2) or its share of future
 * walker traffic (use_keep = 1); among the resident partitions with the
 * fewest queued walkers the one with the lowest value is evicted.  Engines
 * use it when cfg.policy == LT_SCHED_LEARNED and a trained scorer is
 * attached; otherwise they fall back to the plain rules. */
lt_status lt_scorer_create(const lt_partition *p, uint64_t seed, lt_scorer **out);
void lt_scorer_free(lt_scorer *s);
uint64_t lt_scorer_num_params(void);
/* Add the decisions of e's last traced run (record_trace=1) as training
 * samples: with use_next, the label of output 0 is the partition the traced
 * scheduler loaded or read over PCIe at its next decision.  With use_keep =
 * 1 the label of output 1 is each partition's share of the walkers the trace
 * processes after the decision (times P); with use_keep = 2 it is Belady's
 * next-use quantity: with d the number of later records of kind 0, 2 or 3
 * up to and including the partition's next one, 1 - ln(1+d)/ln(1+len), and
 * 0 if it is never needed in HBM again (the eviction of the lowest score is
 * then Belady's choice; cf. lt_sched_belady).  Returns the samples added. */
lt_status lt_scorer_add_trace(lt_scorer *s, const lt_engine *e, int32_t use_next, int32_t use_keep,
                              uint64_t *added);
/* Adam on the accumulated samples; loss_out[2] = last epoch's mean losses. */
lt_status lt_scorer_fit(lt_scorer *s, uint32_t epochs, uint32_t batch, float lr, uint64_t seed, float *loss_out);
/* Explicit samples: n decisions of P partitions each (layouts as in
 * lt_engine_trace_state); keep_label f32[n*P] may be NULL; wa/wb are the
 * per-decision loss weights of the two outputs (NULL = 1 and 0). */
lt_status lt_scorer_train(lt_scorer *s, uint64_t n, const uint32_t *counts, const uint8_t *resident,
                          const uint32_t *age, const int32_t *recent, const int32_t *cur, const float *w_total,
                          const int32_t *next_label, const float *keep_label, const float *wa,
                          const float *wb, uint32_t epochs, uint32_t batch, float lr, uint64_t seed,
                          float *loss_out);
/* One forward + backward without an update: losses (f32[2]) and the
 * gradient w.r.t. the parameters (f32[lt_scorer_num_params()]). */
lt_status lt_scorer_loss_grad(lt_scorer *s, uint64_t n, const uint32_t *counts, const uint8_t *resident,
                              const uint32_t *age, const int32_t *recent, const int32_t *cur, const float *w_total,
                              const int32_t *next_label, const float *keep_label, const float *wa,
                              const float *wb, float *loss_out, float *grad_out);
/* Scores (f32[P*2]) of one decision. */
lt_status lt_scorer_infer(lt_scorer *s, const uint32_t *counts, const uint8_t *resident, const uint32_t *age,
                          const int32_t *recent, int32_t cur, float w_total, float *out);
/* lt_scorer_infer for callers that cannot allocate (code under HIP graph
 * capture, a scheduler's tight decision loop): no counterpart in the
 * reference.  Same scores as lt_scorer_infer, computed through a static
 * per-GPU scratch (about 1.4 MB, held until the process ends) instead of the
 * scorer's staging buffers; calls no allocating routine.  Scorers of at most
 * 65,535 partitions (LT_E_ARG beyond).  Calls on one GPU share its scratch:
 * calls that overlap on one GPU (from several threads) race. */
lt_status lt_scorer_infer_noalloc(lt_scorer *s, const uint32_t *counts, const uint8_t *resident,
                                  const uint32_t *age, const int32_t *recent, int32_t cur, float w_total,
                                  float *out);
/* FP32 master parameters: W1[16][256] b1[256] W2[256][256] b2[256] W3[256][2] b3[2]. */
lt_status lt_scorer_get_params(const lt_scorer *s, float *out);
lt_status lt_scorer_set_params(lt_scorer *s, const float *in);
/* Weights file: "LTMLP2\0\0", u32 16, 256, 2, 0, then the BF16 parameters
 * (files of the earlier feature set, "LTMLP1", are refused). */
lt_status lt_scorer_save(const lt_scorer *s, const char *path);
lt_status lt_scorer_load(lt_scorer *s, const char *path);
/* The scorer the LEARNED policy of e consults (NULL: the plain rules).  The
 * engine keeps the pointer: s must stay alive while e may run with it (set
 * NULL or destroy e first).  A scorer serves one engine at a time (its
 * inference staging is its own).  LT_E_ARG if s was built for another
 * partition count. */
lt_status lt_engine_set_scorer(lt_engine *e, lt_scorer *s);

/* ---- pure scheduler rules (host, no GPU) -------------------------------- */
/* Scheduler::nextPartition greedy (src/Scheduler.h:368-388): argmax of
 * counts over p != prev unless sync; returns -1 when nothing is picked. */
int32_t lt_sched_greedy_next(uint32_t P, const uint64_t *counts, int32_t prev, int32_t sync);
/* GPUScheduler::findPartitionwithleastWalker (src/Scheduler.h:149-171):
 * resident partition != except with the fewest walkers (lowest id on ties). */
int32_t lt_sched_evict_min(uint32_t P, const uint64_t *counts, const uint8_t *resident, int32_t except);
/* GPUScheduler::getComputedPartition (src/Scheduler.h:84-111), greedy variant:
 * a resident partition != except whose queue holds >= page walkers and has the
 * fewest walkers, else the resident one with the most (>0) walkers. */
int32_t lt_sched_computed_partition(uint32_t P, const uint64_t *counts, const uint8_t *resident,
                                    int32_t except, uint64_t page);

/* Belady's MIN on a recorded demand sequence (no counterpart in the
 * reference; the bound for the eviction decision of
 * GPUScheduler::findPartitionwithleastWalker, src/Scheduler.h:149-186): seq[0..n) are
 * the partitions that had to be resident, in order (for an engine trace, the
 * records of kind 0 and 3).  A cache of `slots` slots, holding
 * init_resident's partitions at first (u8[P], NULL = empty; at most `slots`
 * of them, lowest ids first), loads a demanded partition that is absent and,
 * when full, evicts the resident partition whose next demand lies farthest
 * ahead.  *loads_out = its loads, the fewest any eviction rule needs on this
 * sequence; *bytes_out (may be NULL) = the sum of part_bytes (u64[P], may be
 * NULL) over them; victims_out (i32[n], may be NULL) = the partition evicted
 * at each demand, or -1. */
lt_status lt_sched_belady(const int32_t *seq, uint64_t n, uint32_t P, uint32_t slots, const uint8_t *init_resident,
                          const uint64_t *part_bytes, uint64_t *loads_out, uint64_t *bytes_out, int32_t *victims_out);

/* ---- device probe -------------------------------------------------------- */
lt_status lt_device_info(char *name, size_t cap, int *num_cu, uint64_t *hbm_bytes);
/* Device-to-device copy of `bytes` (16-byte vector loads and stores) by the
 * kernel copy_probe_kernel, reps times; gbps_out = bytes read + written per
 * second.  A kernel of known traffic, used to scale HBM traffic counters. */
lt_status lt_device_copy_bandwidth(uint64_t bytes, uint32_t reps, double *gbps_out);
/* Host-to-HBM DMA of `bytes` from pinned host memory (hipMemcpyAsync), reps
 * times; gbps_out = bytes per second.  The achievable rate of the copies that
 * load partitions (GPUGraph::loadAsync, src/GraphLoader.h:193-215), to set
 * beside the link's theoretical rate. */
lt_status lt_device_h2d_bandwidth(uint64_t bytes, uint32_t reps, double *gbps_out);

#ifdef __cplusplus
}
#endif
#endif /* LIGHTTRAFFIC_H */
