// lt_device.h — device-side data layout and kernels of the walk engine.
//
// Walker queues (role of DeviceListPool, src/buffer/Pool.h:372-462): one slab
// of fixed-size chunks of walker records, a per-partition table of chunk ids
// (frame) and a per-partition walker count (qlen).  Chunks are handed out by a
// bump pointer over a free stack while a kernel runs and returned by a
// single-workgroup fix-up kernel between launches, so no host round trip is
// needed to grow a queue (the reference pops pages with <<<1,1>>> kernels and
// reads the page pointers back, Pool.h:616-640).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lt_rng.h"

namespace lt {

constexpr int kBlock = 256;                 // threads per workgroup (4 wave64)
constexpr int kWpt = 4;                     // walkers per thread per pass
constexpr int kBlockWalkers = kBlock * kWpt;  // 1024 walkers per workgroup pass
constexpr int kChunkLog2 = 12;
constexpr uint32_t kChunk = 1u << kChunkLog2;  // 4096 walkers = 64 KiB per chunk
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr uint32_t kMaxPartitions = 65535;  // partition ids are staged as u16 (and .part stores a u16 count)
constexpr uint32_t kSliceChunks = 2048;     // chunks of one source queue per launch
constexpr uint32_t kLostCap = 8192;         // chunks lost to allocation races per launch
constexpr uint32_t kQStride = 32;           // u32 words per queue length: one 128-byte line each,
                                            // so appends to different queues do not share a line
constexpr uint32_t kStatSlots = 256;        // event counters are spread over this many lines
constexpr uint32_t kMaxZcBatch = 16;        // partitions walked from host memory by one launch
constexpr uint32_t kMaxHostPieces = 32;     // spilled walker pieces (<= kChunk each) read by one launch
constexpr uint32_t kCountCopies = 8;        // visit-count copies, one per XCD of an MI355X: workgroups
                                            // are dispatched to the 8 XCDs round-robin, so workgroup b
                                            // adds into copy b % 8 and its lines stay in one XCD's L2

// Walker record: one dwordx4.  The reference's walkers are {current, step}
// (app/pagerank.cu:8-11) or {current, step, id} (app/genericwalk.cu:6-10);
// the id is needed here because every random draw is keyed by it.
struct __attribute__((aligned(16))) Walker {
    uint32_t cur, step, id_lo, id_hi;
};

struct PoolCtl {
    uint32_t top;       // free chunks on the stack at launch start
    uint32_t popped;    // chunks taken during the launch
    uint32_t nlost;     // chunks taken but not linked (lost a CAS)
    uint32_t overflow;  // bit 0: pool exhausted, bit 1: lost list full, bit 2: frame index, bit 3: bad chunk id
};

struct Queues {
    Walker *slab;
    uint32_t *frame;       // [P][maxc] chunk ids, kEmpty when unused
    uint32_t *qlen;        // [P * kQStride] walkers queued in partition p at qlen[p * kQStride]
    uint32_t *free_stack;  // [nchunks]
    PoolCtl *ctl;
    uint32_t *lost;        // [kLostCap]
    uint32_t *packed;      // [P + 4]: the P queue lengths, then PoolCtl (one D2H per decision)
    uint32_t maxc;         // chunk slots per partition
    uint32_t nchunks;
    uint32_t P;
};

__host__ __device__ __forceinline__ uint32_t &qlen_of(const Queues &q, uint32_t p) {
    return q.qlen[(size_t)p * kQStride];
}

struct Parts {
    const uint32_t *node_off;  // [P+1]
    const uint64_t *edge_off;  // [P+1]
    const uint32_t *lut;       // partition of vertex (j << lut_shift)
    const uint64_t *const *rowptr_tab;  // [P] HBM slot rowptr, nullptr if not resident
    const uint32_t *const *col_tab;     // [P] raw col slice, or the block headers of a packed slot
    const uint32_t *const *words_tab;   // [P] bit streams of a packed slot, nullptr for raw columns
    const uint8_t *zc_part;             // [P] 1 while the table points p at host memory (zero-copy batch)
    uint32_t lut_shift;
    uint32_t P;
};

enum AppKind : int { APP_PAGERANK = 0, APP_PPR = 1, APP_GENERIC = 2 };

struct AppArgs {
    uint64_t seed;
    uint32_t V;
    uint32_t thresh;      // pagerank: teleport threshold; ppr: stop threshold
    uint32_t length;      // pagerank/genericwalk: l
    uint32_t source;      // single-source ppr
    uint32_t *counts;     // u32[V] or, with counts_stride = V, u32[kCountCopies][V]; or nullptr
    uint64_t counts_stride;  // 0: one count array; V: workgroup b counts into copy b % kCountCopies
    uint32_t *endpoints;  // u32[N] indexed by walker id, or nullptr
    uint32_t *moves_out;  // u32[N] (genericwalk), or nullptr
    uint32_t *paths;      // genericwalk: u32[N][length + 1], vertex after k moves at [id][k], or nullptr
    uint32_t rp_host;     // 1: zero-copied spans read rowptr over PCIe too; 0: rowptr is in HBM
    const uint64_t *q_off;  // batched ppr: [nq+1] first walker id of each query
    const uint32_t *q_src;  // batched ppr: [nq] sources
    uint32_t nq;
};

// Event counters: stats[slot * kNumStats + StatSlot], summed over kStatSlots slots.
enum StatSlot { ST_MOVES = 0, ST_TELEPORTS, ST_VISITS, ST_RESIDENCIES, ST_ZC_MOVES, ST_ZC_READS, ST_DROPPED, kNumStats = 8 };

// Device-side scheduling (the whole decide / load / walk / fix-up loop runs
// on the GPU, captured in a hipGraph).  One decision per iteration.
struct DevSched {
    int32_t choice;     // partition walked this iteration, -1 = every queue is empty
    int32_t zerocopy;   // walked from host memory
    int32_t load;       // copied into `slot` first
    int32_t slot;
    int32_t victim;     // partition evicted for the load, -1 = none
    uint32_t iterations, loads, zc_calls;
    unsigned long long h2d_bytes;
};

struct SchedState {
    DevSched *ds;             // nullptr: the host schedules
    int32_t *slot_of;         // [P] slot or -1
    int32_t *part_in_slot;    // [nslots] partition or -1
    uint8_t *resident;        // [P]
    const uint64_t **rowptr_tab;  // writable views of Parts' tables
    const uint32_t **col_tab;
    uint8_t *slots;
    uint64_t slot_bytes;
    uint32_t nslots;
    const uint64_t *zc_rowptr;  // rowptr for zero-copied spans: the pinned host graph's, or an HBM copy
    const uint32_t *zc_col;     // device view of the pinned host graph's col
    uint64_t zc_threshold;
    uint32_t rp_host;           // 1: zc_rowptr is host memory (counted as H2D by the copies)
};

// Block slots in the device loop (decide_kernel, load_kernel): partition p's
// block image is size[p] bytes at img + off[p] (mapped host memory), and its
// bit streams are entered in words_tab when it is loaded; img == nullptr or
// size[p] == 0: raw columns.
struct BlockImages {
    const uint32_t **words_tab;  // writable view of Parts::words_tab
    const uint8_t *img;
    const uint64_t *off;
    const uint64_t *size;
};

struct WalkArgs {
    AppArgs app;
    SchedState ss;
    Queues q;
    Parts pt;
    uint32_t src;         // partition whose queue is consumed
    uint32_t chunk_lo;    // first chunk of the slice
    uint32_t chunk_hi;    // one past the last chunk of the slice
    const uint64_t *rowptr;  // rowptr slice of src (HBM slot or mapped host)
    const uint32_t *col;     // col slice of src (block headers if words is set)
    const uint32_t *words;   // bit streams of src's packed slot, nullptr for raw columns
    int zerocopy;         // src slice is read over PCIe
    int span_resident;    // keep walking through HBM-resident partitions
    // How a pass appends its leaving walkers to their queues: 0, one queue
    // reservation per (workgroup, destination) and the walkers staged in LDS
    // so the stores are coalesced; 1, the same reservations with each walker
    // stored from its own lane (the reference's NO_RANDOM_ACCESS_REDUCE
    // build, src/WalkManager.h:98-118); 2, one reservation per walker (its
    // ONE_LEVEL_QUEUE build, src/WalkManager.h:195-209).
    int append_mode;
    unsigned long long *stats;
    // Batch (span_resident only): nsrc > 0 consumes the whole queues of
    // srcs[0..nsrc), each entered in the partition table for the launch
    // (resident slots, or views of the pinned host graph for a zero-copy
    // batch), so a walker moving between them keeps walking and no queue
    // being consumed is appended to.  Queue k occupies [voff[k], voff[k+1])
    // of one virtual range, each start a multiple of kChunk.
    uint32_t nsrc;
    uint32_t srcs[kMaxZcBatch];
    uint64_t voff[kMaxZcBatch + 1];
    // Spilled walkers (nhost > 0): walkers of src kept in pinned host memory,
    // read by the kernel over PCIe; piece k holds hn[k] <= kChunk walkers at
    // hptr[k] and occupies [k * kChunk, k * kChunk + hn[k]) of the range.
    uint32_t nhost;
    const Walker *hptr[kMaxHostPieces];
    uint32_t hn[kMaxHostPieces];
};

// A spill round: the whole device queues of nv partitions are copied to their
// host queues and their chunks returned to the pool.
struct SpillArgs {
    Queues q;
    uint32_t nv;
    const uint32_t *part;    // [nv] partitions spilled
    const uint64_t *prefix;  // [nv + 1] walkers of the partitions before k
    const uint32_t *cprefix; // [nv + 1] device chunks of the partitions before k
    const uint32_t *hstart;  // [nv] slot of partition k's first walker in its first host chunk
    const uint32_t *hbase;   // [nv] index of that host chunk in hptr
    Walker *const *hptr;     // host chunks, mapped for the device
};

// Table entries of a zero-copy batch.
struct ZcBatch {
    uint32_t n;
    uint32_t part[kMaxZcBatch];
    const uint64_t *rowptr[kMaxZcBatch];
    const uint32_t *col[kMaxZcBatch];
};

struct InitArgs {
    AppArgs app;
    Queues q;
    Parts pt;
    uint64_t id_lo, id_hi;  // walker ids created by this launch
    unsigned long long *stats;
};

}  // namespace lt
