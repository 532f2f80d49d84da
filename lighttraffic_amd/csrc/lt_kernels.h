// lt_kernels.h — walk, init and queue kernels (included by lt_engine.hip only).
//
// One walker per lane, kWpt walkers per thread.  A workgroup pass takes 1024
// walker records from the source queue (coalesced dwordx4 loads), walks each
// one while its vertex stays inside the span being processed, then appends
// the walkers that left to their destination partitions' queues:
//   1. LDS histogram of destination partitions (ds atomics, returns the rank
//      of the walker inside its bucket),
//   2. one global atomicAdd per (workgroup, destination) reserves a range of
//      queue slots, and an exclusive scan of the histogram gives each bucket
//      its place in an LDS staging buffer,
//   3. walkers are staged in LDS grouped by destination and written out by
//      consecutive lanes to consecutive queue slots.
// This is the two-level reshuffle of insertToQueueKernel
// (src/WalkManager.h:60-190) fused into the walk kernel
// (randomWalkKernel, src/WalkManager.h:229-246), so a residency costs one
// read and one write of the 16-byte record instead of two of each.
#pragma once
#include "lt_device.h"

namespace lt {

__device__ __forceinline__ uint32_t find_part(const Parts &pt, uint32_t v) {
    uint32_t p = pt.lut[v >> pt.lut_shift];
    while (v >= pt.node_off[p + 1]) ++p;
    return p;
}

__device__ __forceinline__ uint32_t load_agent(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Chunk k of partition part's queue: link a fresh chunk unless another lane
// linked one first (then return ours to the lost list).  Returns kEmpty if the
// pool is exhausted (flagged in ctl->overflow; the host turns it into an error).
__device__ uint32_t alloc_chunk(const Queues &q, uint32_t part, uint32_t k) {
    if (k >= q.maxc) {
        atomicOr(&q.ctl->overflow, 4u);
        return kEmpty;
    }
    uint32_t *slot = &q.frame[(size_t)part * q.maxc + k];
    const uint32_t top = q.ctl->top;
    const uint32_t i = atomicAdd(&q.ctl->popped, 1u);
    if (i >= top) {
        atomicOr(&q.ctl->overflow, 1u);
        return load_agent(slot);
    }
    const uint32_t c = q.free_stack[top - 1 - i];
    const uint32_t old = atomicCAS(slot, kEmpty, c);
    if (old == kEmpty) return c;
    const uint32_t j = atomicAdd(&q.ctl->nlost, 1u);
    if (j < kLostCap) q.lost[j] = c;
    else atomicOr(&q.ctl->overflow, 2u);
    return old;
}

// Chunk k of a queue whose first slot another workgroup reserved.  That
// workgroup took its slot range with an earlier atomicAdd, so it is running
// and links the chunk within a few microseconds: wait for the link (bounded),
// then link it ourselves if it still has not appeared.
__device__ __forceinline__ uint32_t get_chunk(const Queues &q, uint32_t part, uint32_t k) {
    if (k >= q.maxc) {
        atomicOr(&q.ctl->overflow, 4u);
        return kEmpty;
    }
    const uint32_t *slot = &q.frame[(size_t)part * q.maxc + k];
    uint32_t c = load_agent(slot);
    for (int it = 0; c == kEmpty && it < 4096; it++) {
        __builtin_amdgcn_s_sleep(4);
        c = load_agent(slot);
    }
    return c != kEmpty ? c : alloc_chunk(q, part, k);
}

// LDS layout of the append stage for P partitions.
struct AppendLds {
    uint32_t *hist;   // [P] bucket sizes, then the first reserved queue slot
    uint32_t *scan;   // [P+1] exclusive prefix of the bucket sizes
    uint32_t *ch0;    // [P] chunk holding the first reserved slot
    uint32_t *ch1;    // [P] chunk holding the last reserved slot
    Walker *stage;    // [kBlockWalkers] walkers grouped by destination
    uint16_t *dst;    // [kBlockWalkers] destination of each staged walker
};

// walkers staged per workgroup pass = kBlock * wpt
__host__ __device__ inline size_t append_lds_bytes(uint32_t P, int wpt) {
    size_t b = (4 * (size_t)P + 1) * 4;
    b = (b + 15) & ~(size_t)15;
    return b + (size_t)kBlock * wpt * (16 + 2);
}

__device__ __forceinline__ AppendLds carve_lds(uint8_t *smem, uint32_t P, int wpt) {
    AppendLds L;
    L.hist = (uint32_t *)smem;
    L.scan = L.hist + P;
    L.ch0 = L.scan + P + 1;
    L.ch1 = L.ch0 + P;
    size_t off = ((4 * (size_t)P + 1) * 4 + 15) & ~(size_t)15;
    L.stage = (Walker *)(smem + off);
    L.dst = (uint16_t *)(L.stage + kBlock * wpt);
    return L;
}

// Exclusive scan of in[0..n) into out[], returns the total (all threads).
__device__ uint32_t block_exclusive_scan(const uint32_t *in, uint32_t *out, uint32_t n, uint32_t *s_wsum) {
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t per = (n + kBlock - 1) / kBlock;
    const uint32_t beg = min(tid * per, n), end = min(beg + per, n);
    uint32_t sum = 0;
    for (uint32_t i = beg; i < end; i++) sum += in[i];
    uint32_t x = sum;
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) s_wsum[wave] = x;
    __syncthreads();
    uint32_t before = 0, total = 0;
#pragma unroll
    for (uint32_t w = 0; w < kBlock / 64; w++) {
        const uint32_t s = s_wsum[w];
        before += (w < wave) ? s : 0u;
        total += s;
    }
    uint32_t run = before + x - sum;
    for (uint32_t i = beg; i < end; i++) {
        out[i] = run;
        run += in[i];
    }
    return total;
}

// Append the walkers with dst[k] != kEmpty to their partitions' queues.
// Collective: every thread of the workgroup must call it.  Chunk ids are
// resolved once per (workgroup, destination): a reservation of at most
// kBlockWalkers <= kChunk slots touches at most two chunks; the range that
// contains a chunk's first slot links it, and a later range that arrives
// before that link is published links it itself (the CAS loser's chunk goes
// to the lost list).
template <int WPT>
__device__ __forceinline__ void block_append(const Queues &q, uint32_t P, const AppendLds &L, const Walker (&w)[WPT],
                             const uint32_t (&dst)[WPT], uint32_t *s_wsum) {
    const uint32_t tid = threadIdx.x;
    for (uint32_t i = tid; i < P; i += kBlock) L.hist[i] = 0;
    __syncthreads();
    uint32_t rank[WPT];
#pragma unroll
    for (int k = 0; k < WPT; k++)
        if (dst[k] != kEmpty) rank[k] = atomicAdd(&L.hist[dst[k]], 1u);
    __syncthreads();
    const uint32_t total = block_exclusive_scan(L.hist, L.scan, P, s_wsum);
    if (tid == 0) L.scan[P] = total;
    __syncthreads();  // the scan reads hist segments that the loop below overwrites
    // Pass 1: reserve and link the chunks this workgroup owns.  Pass 2 waits
    // for chunks owned by others only after every owned link is published,
    // so no owner ever waits before publishing.
    for (uint32_t i = tid; i < P; i += kBlock) {
        const uint32_t c = L.hist[i];
        if (c) {
            const uint32_t b = atomicAdd(&qlen_of(q, i), c);
            const uint32_t k0 = b >> kChunkLog2, k1 = (b + c - 1) >> kChunkLog2;
            const uint32_t c0 = (b & (kChunk - 1)) == 0 ? alloc_chunk(q, i, k0) : kEmpty;
            L.ch0[i] = c0;
            L.ch1[i] = k1 != k0 ? alloc_chunk(q, i, k1) : c0;
            L.hist[i] = b;
        }
    }
    for (uint32_t i = tid; i < P; i += kBlock) {
        const uint32_t b = L.hist[i];
        if (L.scan[i + 1] != L.scan[i] && (b & (kChunk - 1)) != 0) {
            const uint32_t c0 = get_chunk(q, i, b >> kChunkLog2);
            const uint32_t k1 = (b + (L.scan[i + 1] - L.scan[i]) - 1) >> kChunkLog2;
            L.ch0[i] = c0;
            if (k1 == (b >> kChunkLog2)) L.ch1[i] = c0;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < WPT; k++) {
        if (dst[k] != kEmpty) {
            const uint32_t pos = L.scan[dst[k]] + rank[k];
            L.stage[pos] = w[k];
            L.dst[pos] = (uint16_t)dst[k];
        }
    }
    __syncthreads();
    for (uint32_t i = tid; i < total; i += kBlock) {
        const uint32_t part = L.dst[i];
        const uint32_t b = L.hist[part];
        const uint32_t s = b + (i - L.scan[part]);
        const uint32_t c = (s >> kChunkLog2) == (b >> kChunkLog2) ? L.ch0[part] : L.ch1[part];
        if (c < q.nchunks) q.slab[(size_t)c * kChunk + (s & (kChunk - 1))] = L.stage[i];
        else atomicOr(&q.ctl->overflow, 8u);
    }
    __syncthreads();
}

// Per-thread event counts of one launch (far below 2^32), summed as u64.
struct Counters {
    uint32_t moves = 0, tele = 0, visits = 0, resid = 0, zc_moves = 0, zc_reads = 0;
    uint32_t *vkeys = nullptr, *vcnt = nullptr;  // workgroup visit table in LDS (walk kernels)
};

// Workgroup sum of the counters, added to slot blockIdx.x % kStatSlots, so
// that workgroups of one launch do not all update one cache line.
__device__ void flush_counters(const Counters &c, unsigned long long *stats) {
    __shared__ unsigned long long s_part[kBlock / 64][6];
    const uint32_t v[6] = {c.moves, c.tele, c.visits, c.resid, c.zc_moves, c.zc_reads};
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < 6; i++) {
        unsigned long long x = v[i];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        if (lane == 0) s_part[wave][i] = x;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        unsigned long long x = 0;
#pragma unroll
        for (int w = 0; w < kBlock / 64; w++) x += s_part[w][threadIdx.x];
        if (x) atomicAdd(&stats[(size_t)(blockIdx.x % kStatSlots) * kNumStats + threadIdx.x], x);
    }
}

// The graph span a walker is currently walking in.
struct Span {
    const uint64_t *rowptr;  // rowptr[v - nb] = global offset of v's first edge
    const uint32_t *col;     // col[e - eb], or the block headers of a packed slot
    const uint32_t *words;   // bit streams of a packed slot, nullptr for raw columns
    uint64_t eb;
    uint32_t nb, ne;
    int zc;
};

__device__ __forceinline__ uint64_t walker_id(const Walker &w) { return ((uint64_t)w.id_hi << 32) | w.id_lo; }

// Visits go to a workgroup table in LDS (open addressing, 4 probes) and reach
// the global counters once per vertex and flush (flush_visits, at every
// append of staged walkers and at the end); a vertex that finds no slot is
// counted in global memory directly.
constexpr uint32_t kVisitSlots = 2048;
constexpr uint32_t kVisitEmpty = 0xFFFFFFFFu;

__device__ __forceinline__ void count_visit(const AppArgs &a, uint32_t v, Counters &ct) {
    if (ct.vkeys) {
        const uint32_t h = (v * 2654435761u) >> 21;  // 11 bits: kVisitSlots
#pragma unroll
        for (uint32_t probe = 0; probe < 4; probe++) {
            const uint32_t i = (h + probe) & (kVisitSlots - 1);
            uint32_t k = ct.vkeys[i];
            if (k == kVisitEmpty) k = atomicCAS(&ct.vkeys[i], kVisitEmpty, v);
            if (k == kVisitEmpty || k == v) {
                atomicAdd(&ct.vcnt[i], 1u);
                return;
            }
        }
    }
    atomicAdd(&a.counts[(size_t)(blockIdx.x % kCountCopies) * a.counts_stride + v], 1u);
}

// Adds the workgroup's visit table to the global counters and empties it.
// Collective over the workgroup.
__device__ __forceinline__ void flush_visits(const AppArgs &a, const Counters &ct) {
    if (!ct.vkeys) return;
    __syncthreads();
    uint32_t *g = a.counts + (size_t)(blockIdx.x % kCountCopies) * a.counts_stride;
    for (uint32_t i = threadIdx.x; i < kVisitSlots; i += blockDim.x) {
        const uint32_t k = ct.vkeys[i];
        if (k != kVisitEmpty) {
            atomicAdd(&g[k], ct.vcnt[i]);
            ct.vkeys[i] = kVisitEmpty;
            ct.vcnt[i] = 0;
        }
    }
    __syncthreads();
}

__device__ __forceinline__ void finish_walker(const Walker &w, const AppArgs &a) {
    if (a.endpoints) a.endpoints[walker_id(w)] = w.cur;
}

// Walker steps are split into stages so that the K walkers of a thread step
// in lockstep and their dependent reads overlap: stage_a draws the step's
// Philox words (and counts the visit), stage_b reads the rowptr pair of
// every walker, stage_c the chosen column (or, in a packed slot, its block
// header), stage_d the bit-stream words of packed slots, and stage_e moves
// the walkers.  Per walker the draws, reads and updates are exactly those of
// one App::update; only their order across walkers changes.
struct StepTmp {
    u32x4 x;        // Philox words of the step
    uint64_t b, d;  // first edge (global) and degree of the current vertex
    uint64_t e;     // partition-local edge taken; packed slot: bit position of its value
    uint2 h;        // packed slot: block header {base, bit offset}
    uint32_t nx;    // packed slot: next block's bit offset, then the value width
    uint32_t c, c2; // column (raw); packed slot: the two stream words holding the value
    bool need;      // reads the rowptr pair this step
    bool mv;        // moves to a neighbour this step
};

// Stage a of app/pagerank.cu:26-44 (count the visit; after the last step
// nothing moves), app/ppr.cu:22-42 (count the endpoint when one step is
// left) and app/genericwalk.cu:22-32 (record the path).  A walker whose walk
// just ended keeps need = false and step = 0.
template <int APP>
__device__ __forceinline__ void stage_a(Walker &w, const AppArgs &a, Counters &ct, StepTmp &t) {
    t.mv = false;
    t.need = false;
    if (APP == APP_PAGERANK) {
        count_visit(a, w.cur, ct);
        ct.visits++;
        w.step -= 1;
        if (w.step == 0) return;
        t.x = draw(a.seed, walker_id(w), w.step, TAG_PAGERANK);
        t.need = !(t.x.x < a.thresh);  // teleport without reading the graph
    } else if (APP == APP_PPR) {
        if (w.step == 1) {
            if (a.counts) count_visit(a, w.cur, ct);
            ct.visits++;
        }
        w.step -= 1;
        if (w.step == 0) return;
        t.x = draw(a.seed, walker_id(w), w.step, TAG_PPR);
        t.need = true;
    } else {
        if (a.paths) a.paths[walker_id(w) * ((uint64_t)a.length + 1) + (a.length - w.step)] = w.cur;
        t.x = draw(a.seed, walker_id(w), w.step, TAG_GENERIC);
        t.need = true;
    }
}

__device__ __forceinline__ void stage_b(const Walker &w, const Span &sp, StepTmp &t) {
    if (!t.need) return;
    const uint64_t *r = sp.rowptr + (w.cur - sp.nb);
    t.b = r[0];
    t.d = r[1];
}

// Column of partition-local edge e.  A packed slot holds, per block of 16
// edges, {base, bit offset} (the next block's offset gives the width) and
// the bit streams: column = base + bits(offset + (e % 16) * width, width).
__device__ __forceinline__ void stage_c(const Span &sp, StepTmp &t) {
    if (!t.need) return;
    t.d -= t.b;
    if (t.d == 0) return;
    t.mv = true;
    t.e = t.b - sp.eb + bounded(t.x.y, t.x.z, t.d);
    if (!sp.words) {
        t.c = sp.col[t.e];
    } else {
        const uint64_t k = t.e >> 4;
        t.h = ((const uint2 *)sp.col)[k];
        t.nx = sp.col[2 * k + 3];
    }
}

__device__ __forceinline__ void stage_d(const Span &sp, StepTmp &t) {
    if (!t.mv || !sp.words) return;
    const uint32_t w = (t.nx - t.h.y) >> 4;
    const uint32_t pos = t.h.y + (uint32_t)(t.e & 15) * w;
    t.c = sp.words[pos >> 5];
    t.c2 = sp.words[(pos >> 5) + 1];
    t.nx = w;
    t.e = pos;
}

__device__ __forceinline__ uint32_t step_col(const Span &sp, const StepTmp &t) {
    if (!sp.words) return t.c;
    const uint32_t bits = __builtin_amdgcn_alignbit(t.c2, t.c, (uint32_t)t.e & 31u);
    return t.h.x + (t.nx >= 32 ? bits : bits & ((1u << t.nx) - 1u));
}

// Stage e: pagerank teleports when it drew a jump or stands on a vertex
// without neighbours (the reference divides by zero there); ppr counts a
// vertex without neighbours and stops there; genericwalk stops there and
// records its number of moves.
template <int APP>
__device__ __forceinline__ void stage_e(Walker &w, const Span &sp, const AppArgs &a, Counters &ct, const StepTmp &t) {
    if (t.mv) {
        w.cur = step_col(sp, t);
        ct.moves++;
        if (sp.zc) {
            ct.zc_moves++;
            ct.zc_reads += 1 + a.rp_host;
        }
        if (APP == APP_GENERIC) {
            w.step -= 1;
            if (w.step == 0) {
                if (a.moves_out) a.moves_out[walker_id(w)] = a.length;
                if (a.paths) a.paths[walker_id(w) * ((uint64_t)a.length + 1) + a.length] = w.cur;
            }
        }
        return;
    }
    if (t.need && sp.zc) ct.zc_reads += a.rp_host;  // read a rowptr pair of degree 0
    if (APP == APP_PAGERANK) {
        if (w.step == 0) return;
        w.cur = (uint32_t)bounded(t.x.y, t.x.z, a.V);
        ct.tele++;
    } else if (APP == APP_PPR) {
        if (w.step == 0) return;
        if (a.counts) count_visit(a, w.cur, ct);
        ct.visits++;
        w.step = 0;
    } else {
        if (a.moves_out) a.moves_out[walker_id(w)] = a.length - w.step;
        w.step = 0;
    }
}

// Before each lockstep step: end the walk, or find the span the walker's
// vertex lies in.  Returns true while the walker takes steps in this launch;
// otherwise dst is its destination partition, or kEmpty when its walk is
// over.  With span_resident the walker continues into any partition whose
// slot is resident (and into the source slice itself).
__device__ __forceinline__ bool settle(Walker &w, const WalkArgs &a, uint32_t src, const Span &src_sp, Span &sp,
                                       uint32_t &dst) {
    while (true) {
        if (w.step == 0) {
            finish_walker(w, a.app);
            dst = kEmpty;
            return false;
        }
        if ((w.cur - sp.nb) < (sp.ne - sp.nb)) return true;
        const uint32_t p = find_part(a.pt, w.cur);
        if (!a.span_resident) {
            dst = p;
            return false;
        }
        if (p == src) {
            sp = src_sp;
            continue;
        }
        const uint64_t *rp = a.pt.rowptr_tab[p];
        if (!rp) {
            dst = p;
            return false;
        }
        sp.rowptr = rp;
        sp.col = a.pt.col_tab[p];
        sp.words = a.pt.words_tab[p];
        sp.nb = a.pt.node_off[p];
        sp.ne = a.pt.node_off[p + 1];
        sp.eb = a.pt.edge_off[p];
        sp.zc = a.pt.zc_part[p];
    }
}

// One update of the walker: the stages in order.
template <int APP>
__device__ __forceinline__ void app_step(Walker &w, const Span &sp, const AppArgs &a, Counters &ct) {
    StepTmp t;
    stage_a<APP>(w, a, ct, t);
    stage_b(w, sp, t);
    stage_c(sp, t);
    stage_d(sp, t);
    stage_e<APP>(w, sp, a, ct, t);
}

// Where the walkers of one launch come from, kept in LDS: the launch's
// walkers form one dense range, and source k holds [pref[k], pref[k+1]).  A
// source is a queue (walkers from queue slot first[k] of partition part[k]
// on) or a piece of spilled walkers in host memory (host[k]).
constexpr uint32_t kMaxSources = kMaxHostPieces > kMaxZcBatch ? kMaxHostPieces : kMaxZcBatch;
struct SrcTable {
    uint64_t pref[kMaxSources + 1];
    uint64_t first[kMaxSources];
    const Walker *host[kMaxSources];
    uint32_t part[kMaxSources];
    uint32_t n;
    uint32_t cursor;  // walkers of this workgroup's share taken so far
    uint32_t staged;  // walkers waiting in the LDS stage for their append
};

// Thread 0 fills the table; the caller synchronises.  src_out is the
// partition whose queue the launch consumes (kEmpty for a zero-copy batch,
// whose partitions are entered in the partition table for the launch).
__device__ void build_sources(const WalkArgs &a, SrcTable &s, uint32_t &src_out) {
    s.n = 0;
    s.pref[0] = 0;
    s.cursor = 0;
    s.staged = 0;
    src_out = a.src;
    if (a.nsrc) {  // zero-copy batch: whole queues of srcs[0..nsrc)
        src_out = kEmpty;
#pragma unroll  // constant indices keep the argument arrays out of scratch memory
        for (uint32_t k = 0; k < kMaxZcBatch; k++) {
            if (k < a.nsrc) {
                s.part[k] = a.srcs[k];
                s.first[k] = 0;
                s.host[k] = nullptr;
                s.pref[k + 1] = s.pref[k] + qlen_of(a.q, a.srcs[k]);
            }
        }
        s.n = a.nsrc;
        return;
    }
    if (a.nhost) {  // walkers of src spilled to host memory
#pragma unroll
        for (uint32_t k = 0; k < kMaxHostPieces; k++) {
            if (k < a.nhost) {
                s.part[k] = a.src;
                s.first[k] = 0;
                s.host[k] = a.hptr[k];
                s.pref[k + 1] = s.pref[k] + a.hn[k];
            }
        }
        s.n = a.nhost;
        return;
    }
    uint32_t src = a.src, chunk_lo = a.chunk_lo, chunk_hi = a.chunk_hi;
    if (a.ss.ds) {  // the decision of this iteration was taken on the device
        if (a.ss.ds->choice < 0) return;
        src = (uint32_t)a.ss.ds->choice;
        chunk_lo = 0;
        chunk_hi = 0xFFFFFFFFu;
    }
    src_out = src;
    const uint64_t lo = (uint64_t)chunk_lo << kChunkLog2;
    const uint64_t hi = min((uint64_t)qlen_of(a.q, src), (uint64_t)chunk_hi << kChunkLog2);
    s.part[0] = src;
    s.first[0] = lo;
    s.host[0] = nullptr;
    s.pref[1] = hi > lo ? hi - lo : 0;
    s.n = 1;
}

// Walker i of the launch's dense range, and the span it starts in.  Returns
// false (and flags the pool) if its queue chunk is not a valid chunk id.
__device__ __forceinline__ bool fetch_walker(const WalkArgs &a, const SrcTable &s, const Span &src_sp, uint64_t i,
                                             Walker &w, Span &sp) {
    uint32_t k = 0;
    while (k + 1 < s.n && i >= s.pref[k + 1]) k++;
    const uint64_t local = i - s.pref[k];
    if (s.host[k]) {
        w = s.host[k][local];
        sp = src_sp;
        return true;
    }
    const uint32_t p = s.part[k];
    const uint64_t pos = s.first[k] + local;
    const uint32_t c = a.q.frame[(size_t)p * a.q.maxc + (uint32_t)(pos >> kChunkLog2)];
    if (c >= a.q.nchunks) {
        atomicOr(&a.q.ctl->overflow, 8u);
        return false;
    }
    w = a.q.slab[(size_t)c * kChunk + (uint32_t)(pos & (kChunk - 1))];
    if (a.nsrc) {  // a zero-copy batch partition, read from the pinned host graph
        sp.rowptr = a.ss.zc_rowptr + a.pt.node_off[p];
        sp.col = a.ss.zc_col + a.pt.edge_off[p];
        sp.words = nullptr;
        sp.nb = a.pt.node_off[p];
        sp.ne = a.pt.node_off[p + 1];
        sp.eb = a.pt.edge_off[p];
        sp.zc = 1;
    } else {
        sp = src_sp;
    }
    return true;
}

// Walk kernel: consumes the walkers of one launch (a slice of a partition's
// queue, a zero-copy batch of whole queues, or spilled pieces in host
// memory).  Each workgroup takes an equal share of the launch's walkers and
// streams it through its lanes: every iteration each lane takes one step of
// its walker; a lane whose walker ended or left the span stages it in LDS
// and takes the next walker of the share at once.  So a lane idles only at
// the end of the share, not while the longest walker of a fixed batch
// finishes.  Once a workgroup's stage holds a workgroup's worth of walkers
// they are appended to their destination queues together (block_append:
// LDS histogram, scan, one atomic per destination).
template <int APP>
__global__ __launch_bounds__(kBlock) void walk_kernel(WalkArgs a) {
    extern __shared__ __attribute__((aligned(16))) uint8_t smem[];
    __shared__ uint32_t s_wsum[kBlock / 64];
    __shared__ SrcTable s_src;
    __shared__ uint32_t s_srcpart;
    constexpr int kStageWpt = 2;  // stage holds kBlock * 2 walkers
    const AppendLds L = carve_lds(smem, a.pt.P, kStageWpt);
    Counters ct;
    if constexpr (APP != APP_GENERIC) {
        __shared__ uint32_t s_vkeys[kVisitSlots], s_vcnt[kVisitSlots];
        if (a.app.counts) {  // uniform
            for (uint32_t i = threadIdx.x; i < kVisitSlots; i += blockDim.x) {
                s_vkeys[i] = kVisitEmpty;
                s_vcnt[i] = 0;
            }
            ct.vkeys = s_vkeys;
            ct.vcnt = s_vcnt;
        }
    }
    if (threadIdx.x == 0) build_sources(a, s_src, s_srcpart);
    __syncthreads();
    const uint64_t total = s_src.pref[s_src.n];
    if (total == 0) return;  // uniform: nothing counted, nothing staged
    const uint32_t src = s_srcpart;
    Span src_sp;
    src_sp.rowptr = a.rowptr;
    src_sp.col = a.col;
    src_sp.words = a.words;
    src_sp.zc = a.zerocopy;
    if (a.ss.ds && src != kEmpty) {  // device-scheduled iteration: raw columns in a slot or in host memory
        src_sp.zc = a.ss.ds->zerocopy;
        src_sp.words = nullptr;
        if (src_sp.zc) {
            src_sp.rowptr = a.ss.zc_rowptr + a.pt.node_off[src];
            src_sp.col = a.ss.zc_col + a.pt.edge_off[src];
        } else {
            const uint8_t *base = a.ss.slots + (size_t)a.ss.slot_of[src] * a.ss.slot_bytes;
            src_sp.rowptr = (const uint64_t *)base;
            src_sp.col = (const uint32_t *)(base + ((uint64_t)a.pt.node_off[src + 1] - a.pt.node_off[src] + 1) * 8);
        }
    }
    if (src != kEmpty) {
        src_sp.nb = a.pt.node_off[src];
        src_sp.ne = a.pt.node_off[src + 1];
        src_sp.eb = a.pt.edge_off[src];
    }
    // this workgroup's share [lo, hi) of the dense range, 64-aligned
    const uint64_t per = ((total + gridDim.x - 1) / gridDim.x + 63) & ~(uint64_t)63;
    const uint64_t lo = min(total, (uint64_t)blockIdx.x * per), hi = min(total, lo + per);
    const uint32_t lane = threadIdx.x & 63u;
    Walker w;
    Span sp = src_sp;
    bool live = false;
    while (true) {
        // lanes without a walker take the next ones of the share, one LDS
        // atomic per wave, consecutive lanes getting consecutive walkers
        const uint64_t want = __ballot(!live);
        if (want) {
            const uint32_t leader = (uint32_t)__ffsll((unsigned long long)want) - 1;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(&s_src.cursor, (uint32_t)__popcll(want));
            base = __shfl(base, leader, 64);
            if (!live) {
                const uint64_t i = lo + base + (uint32_t)__popcll(want & ((1ull << lane) - 1));
                if (i < hi && fetch_walker(a, s_src, src_sp, i, w, sp)) {
                    live = true;
                    ct.resid++;
                }
            }
        }
        if (live) {
            uint32_t dst;
            live = settle(w, a, src, src_sp, sp, dst);
            if (!live && dst != kEmpty) {  // at most one per lane and iteration
                const uint32_t slot = atomicAdd(&s_src.staged, 1u);
                L.stage[slot] = w;
                L.dst[slot] = (uint16_t)dst;
            }
        }
        if (live) app_step<APP>(w, sp, a.app, ct);
        const bool any_live = __syncthreads_or(live) != 0;
        const uint32_t staged = s_src.staged;
        const bool done = !any_live && lo + s_src.cursor >= hi;
        if (staged >= kBlock || (done && staged > 0)) {  // staged < 2 * kBlock: at most kBlock added per iteration
            Walker sw[kStageWpt];
            uint32_t sd[kStageWpt];
#pragma unroll
            for (int j = 0; j < kStageWpt; j++) {
                const uint32_t i = (uint32_t)j * kBlock + threadIdx.x;  // < 2 * kBlock, inside the stage
                sw[j] = L.stage[i];
                sd[j] = i < staged ? (uint32_t)L.dst[i] : kEmpty;
            }
            __syncthreads();  // the stage is read; block_append reuses it
            if (threadIdx.x == 0) s_src.staged = 0;
            block_append<kStageWpt>(a.q, a.pt.P, L, sw, sd, s_wsum);
            flush_visits(a.app, ct);  // keep the visit table for the walkers to come
        }
        if (done) break;
    }
    flush_visits(a.app, ct);
    flush_counters(ct, a.stats);
}

// PPR length (src/walk/SingleSource.h:12-17): 2 + number of draws that do not
// stop before the first one that does.
__device__ __forceinline__ uint32_t ppr_length(uint64_t seed, uint64_t id, uint32_t thresh) {
    uint32_t F = 0;
    for (uint32_t j = 0;; j++) {
        const u32x4 x = draw(seed, id, j, TAG_PPRLEN);
        if (x.x < thresh) return 2u + F;
        if (x.y < thresh) return 3u + F;
        if (x.z < thresh) return 4u + F;
        if (x.w < thresh) return 5u + F;
        F += 4;
        if (F > 0xFFFFFF00u) return 2u + F;
    }
}

// Create walkers [id_lo, id_hi) and append them to their start partitions'
// queues (GPUInitWalker::create + insertPushedToQueue, src/walk/InitWalk.h:16-20,
// src/WalkManager.h:341-350).
template <int APP>
__global__ __launch_bounds__(kBlock) void init_kernel(InitArgs a) {
    extern __shared__ __attribute__((aligned(16))) uint8_t smem[];
    __shared__ uint32_t s_wsum[kBlock / 64];
    const AppendLds L = carve_lds(smem, a.pt.P, kWpt);
    for (uint64_t base = a.id_lo + (uint64_t)blockIdx.x * kBlockWalkers; base < a.id_hi;
         base += (uint64_t)gridDim.x * kBlockWalkers) {
        Walker w[kWpt];
        uint32_t dst[kWpt];
#pragma unroll
        for (int k = 0; k < kWpt; k++) {
            const uint64_t id = base + (uint64_t)k * kBlock + threadIdx.x;
            dst[k] = kEmpty;
            w[k].id_lo = (uint32_t)id;
            w[k].id_hi = (uint32_t)(id >> 32);
            if (id >= a.id_hi) {
                w[k].cur = 0;
                w[k].step = 0;
                continue;
            }
            if (APP == APP_PPR) {
                uint32_t src = a.app.source;
                if (a.app.q_off) {
                    uint32_t lo = 0, hi = a.app.nq;  // last query with q_off[q] <= id
                    while (hi - lo > 1) {
                        const uint32_t mid = (lo + hi) >> 1;
                        if (a.app.q_off[mid] <= id) lo = mid;
                        else hi = mid;
                    }
                    src = a.app.q_src[lo];
                }
                w[k].cur = src;
                w[k].step = ppr_length(a.app.seed, id, a.app.thresh);
            } else {
                const u32x4 x = draw(a.app.seed, id, 0, TAG_START);
                w[k].cur = (uint32_t)bounded(x.x, x.y, a.app.V);
                w[k].step = a.app.length;
            }
            if (w[k].step > 0) {
                dst[k] = find_part(a.pt, w[k].cur);
            } else {
                if (APP == APP_GENERIC && a.app.moves_out) a.app.moves_out[id] = 0;
                if (APP == APP_GENERIC && a.app.paths) a.app.paths[id * ((uint64_t)a.app.length + 1)] = w[k].cur;
                finish_walker(w[k], a.app);
            }
        }
        block_append<kWpt>(a.q, a.pt.P, L, w, dst, s_wsum);
    }
}

// Copy the queue lengths and PoolCtl into q.packed for the host's one D2H.
// Collective over one workgroup; runs after every queue update of the stream.
__device__ void pack_meta(const Queues &q) {
    for (uint32_t i = threadIdx.x; i < q.P; i += blockDim.x) q.packed[i] = qlen_of(q, i);
    if (threadIdx.x == 0) {
        const PoolCtl c = *q.ctl;
        q.packed[q.P] = c.top;
        q.packed[q.P + 1] = c.popped;
        q.packed[q.P + 2] = c.nlost;
        q.packed[q.P + 3] = c.overflow;
    }
}

__global__ __launch_bounds__(kBlock) void pack_meta_kernel(Queues q) { pack_meta(q); }

// Return the popped-but-unused state of the pool, the lost chunks and the
// consumed source chunks [chunk_lo, chunk_hi) to the free stack; clear the
// source queue after its last slice; pack the queue lengths.  One workgroup.
__global__ __launch_bounds__(kBlock) void fixup_kernel(Queues q, uint32_t src, uint32_t chunk_lo, uint32_t chunk_hi,
                                                       int last, const DevSched *ds) {
    __shared__ uint32_t s_top, s_nlost, s_cnt;
    if (ds) {  // device-scheduled iteration: the whole queue of the choice was consumed
        src = ds->choice < 0 ? kEmpty : (uint32_t)ds->choice;
        chunk_lo = 0;
        chunk_hi = 0xFFFFFFFFu;
        last = 1;
    }
    if (threadIdx.x == 0) {
        PoolCtl c = *q.ctl;
        uint32_t popped = c.popped;
        if (popped > c.top) popped = c.top;
        uint32_t nlost = c.nlost < kLostCap ? c.nlost : kLostCap;
        uint32_t cnt = 0;
        if (src != kEmpty) {
            const uint32_t nch = (qlen_of(q, src) + kChunk - 1) >> kChunkLog2;
            const uint32_t hi = min(chunk_hi, nch);
            cnt = hi > chunk_lo ? hi - chunk_lo : 0;
        }
        s_top = c.top - popped;
        s_nlost = nlost;
        s_cnt = cnt;
        if (s_top + nlost + cnt > q.nchunks) {
            q.ctl->overflow |= 16u;
            s_nlost = 0;
            s_cnt = 0;
        }
    }
    __syncthreads();
    const uint32_t top = s_top, nlost = s_nlost, cnt = s_cnt;
    for (uint32_t i = threadIdx.x; i < nlost; i += kBlock) q.free_stack[top + i] = q.lost[i];
    for (uint32_t j = threadIdx.x; j < cnt; j += kBlock) {
        uint32_t *f = &q.frame[(size_t)src * q.maxc + chunk_lo + j];
        q.free_stack[top + nlost + j] = *f;
        *f = kEmpty;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        q.ctl->top = top + nlost + cnt;
        q.ctl->popped = 0;
        q.ctl->nlost = 0;
        if (last && src != kEmpty) qlen_of(q, src) = 0;
    }
    __syncthreads();
    pack_meta(q);
}

// Copy the queues of a spill round into host memory: walker i of partition
// k goes to slot hstart[k] + i of its host chunks (vector stores over PCIe).
__global__ __launch_bounds__(256) void spill_gather_kernel(SpillArgs s) {
    const uint64_t total = s.prefix[s.nv];
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += stride) {
        uint32_t lo = 0, hi = s.nv;  // last k with prefix[k] <= idx
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (s.prefix[mid] <= idx) lo = mid;
            else hi = mid;
        }
        const uint64_t i = idx - s.prefix[lo];
        const uint32_t c = s.q.frame[(size_t)s.part[lo] * s.q.maxc + (uint32_t)(i >> kChunkLog2)];
        if (c >= s.q.nchunks) {
            atomicOr(&s.q.ctl->overflow, 8u);
            continue;
        }
        const uint64_t hp = s.hstart[lo] + i;
        s.hptr[s.hbase[lo] + (uint32_t)(hp >> kChunkLog2)][hp & (kChunk - 1)] =
            s.q.slab[(size_t)c * kChunk + (i & (kChunk - 1))];
    }
}

// Return the spilled partitions' chunks to the free stack and empty their
// queues; pack the queue lengths.  One workgroup, after spill_gather_kernel.
__global__ __launch_bounds__(kBlock) void release_queues_kernel(SpillArgs s) {
    __shared__ uint32_t s_top;
    if (threadIdx.x == 0) {
        s_top = s.q.ctl->top;
        if (s_top + s.cprefix[s.nv] > s.q.nchunks) {
            s.q.ctl->overflow |= 16u;
            s_top = 0xFFFFFFFFu;
        }
    }
    __syncthreads();
    const uint32_t top = s_top;
    if (top != 0xFFFFFFFFu) {
        for (uint32_t k = 0; k < s.nv; k++) {
            const uint32_t nch = s.cprefix[k + 1] - s.cprefix[k];
            uint32_t *f = s.q.frame + (size_t)s.part[k] * s.q.maxc;
            for (uint32_t j = threadIdx.x; j < nch; j += blockDim.x) {
                s.q.free_stack[top + s.cprefix[k] + j] = f[j];
                f[j] = kEmpty;
            }
        }
        for (uint32_t k = threadIdx.x; k < s.nv; k += blockDim.x) qlen_of(s.q, s.part[k]) = 0;
    }
    __syncthreads();
    if (threadIdx.x == 0 && top != 0xFFFFFFFFu) s.q.ctl->top = top + s.cprefix[s.nv];
    __syncthreads();
    pack_meta(s.q);
}

// Empty every queue and put all chunks on the free stack.
__global__ void reset_queues_kernel(Queues q, uint32_t P) {
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = tid; i < (size_t)P * q.maxc; i += stride) q.frame[i] = kEmpty;
    for (size_t i = tid; i < q.nchunks; i += stride) q.free_stack[i] = (uint32_t)i;
    for (size_t i = tid; i < (size_t)P * kQStride; i += stride) q.qlen[i] = 0;
    if (tid == 0) {
        q.ctl->top = q.nchunks;
        q.ctl->popped = 0;
        q.ctl->nlost = 0;
        q.ctl->overflow = 0;
    }
}

// Enter (on = 1) or remove a zero-copy batch's partitions in the partition table.
__global__ void set_zc_tables_kernel(const uint64_t **rowptr_tab, const uint32_t **col_tab, uint8_t *zc_part,
                                     ZcBatch b, int on) {
    const uint32_t k = threadIdx.x;
    if (blockIdx.x == 0 && k < b.n) {
        const uint32_t p = b.part[k];
        rowptr_tab[p] = on ? b.rowptr[k] : nullptr;
        col_tab[p] = on ? b.col[k] : nullptr;
        zc_part[p] = on ? 1 : 0;
    }
}

__global__ void set_table_kernel(const uint64_t **rowptr_tab, const uint32_t **col_tab, const uint32_t **words_tab,
                                 uint32_t p, const uint64_t *rp, const uint32_t *cp, const uint32_t *wp) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        rowptr_tab[p] = rp;
        col_tab[p] = cp;
        words_tab[p] = wp;
    }
}

// Plain streaming copy of n 16-byte vectors (known HBM traffic: 16n read, 16n written).
__global__ __launch_bounds__(256) void copy_probe_kernel(const uint4 *src, uint4 *dst, uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// counts[v] = sum over the kCountCopies copies of v's count (in place into copy 0).
__global__ __launch_bounds__(256) void sum_count_copies_kernel(uint32_t *counts, uint32_t V) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += stride) {
        uint32_t s = 0;
#pragma unroll
        for (uint32_t c = 0; c < kCountCopies; c++) s += counts[c * (size_t)V + v];
        counts[v] = s;
    }
}

__global__ void fill_u32_kernel(uint32_t *p, size_t n, uint32_t v) {
    const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (size_t i = tid; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// ---- device-side scheduling ------------------------------------------------
// One workgroup decides the iteration with the plain rules: the partition with
// the most queued walkers (lowest id on ties); walk it in place if resident,
// read it over PCIe if it holds fewer than zc_threshold walkers, else copy it
// into a free slot or into the slot of the resident partition with the fewest
// walkers.  Tables are updated here; the copy is done by load_kernel next.
__global__ __launch_bounds__(1024) void decide_kernel(Queues q, SchedState ss, Parts pt) {
    __shared__ unsigned long long s_best[1024];
    __shared__ unsigned long long s_victim[1024];
    __shared__ uint32_t s_free[1024];
    const uint32_t tid = threadIdx.x, P = pt.P;
    unsigned long long best = 0, victim = ~0ull;
    for (uint32_t p = tid; p < P; p += 1024) {
        const unsigned long long c = qlen_of(q, p);
        const unsigned long long key = (c << 32) | (0xFFFFFFFFull - p);
        if (c > 0 && key > best) best = key;
    }
    s_best[tid] = best;
    __syncthreads();
    for (uint32_t off = 512; off > 0; off >>= 1) {
        if (tid < off && s_best[tid + off] > s_best[tid]) s_best[tid] = s_best[tid + off];
        __syncthreads();
    }
    best = s_best[0];
    DevSched *ds = ss.ds;
    if (best == 0) {
        if (tid == 0) {
            ds->choice = -1;
            ds->load = 0;
            ds->zerocopy = 0;
            ds->victim = -1;
        }
        return;
    }
    const uint32_t c = 0xFFFFFFFFu - (uint32_t)(best & 0xFFFFFFFFull);
    const uint64_t cnt = best >> 32;
    const bool need_slot = !ss.resident[c] && cnt >= ss.zc_threshold;
    uint32_t fr = 0xFFFFFFFFu;
    if (need_slot) {
        for (uint32_t s = tid; s < ss.nslots; s += 1024)
            if (ss.part_in_slot[s] < 0 && s < fr) fr = s;
        for (uint32_t p = tid; p < P; p += 1024) {
            if (!ss.resident[p] || p == c) continue;
            const unsigned long long key = ((unsigned long long)qlen_of(q, p) << 32) | p;
            if (key < victim) victim = key;
        }
    }
    s_free[tid] = fr;
    s_victim[tid] = victim;
    __syncthreads();
    for (uint32_t off = 512; off > 0; off >>= 1) {
        if (tid < off) {
            if (s_free[tid + off] < s_free[tid]) s_free[tid] = s_free[tid + off];
            if (s_victim[tid + off] < s_victim[tid]) s_victim[tid] = s_victim[tid + off];
        }
        __syncthreads();
    }
    if (tid != 0) return;
    ds->choice = (int32_t)c;
    ds->iterations++;
    ds->victim = -1;
    ds->load = 0;
    ds->zerocopy = 0;
    if (ss.resident[c]) return;
    if (!need_slot) {
        ds->zerocopy = 1;
        ds->zc_calls++;
        return;
    }
    uint32_t slot = s_free[0];
    if (slot == 0xFFFFFFFFu) {
        if (s_victim[0] == ~0ull) {  // nothing to evict: read over PCIe instead
            ds->zerocopy = 1;
            ds->zc_calls++;
            return;
        }
        const uint32_t v = (uint32_t)(s_victim[0] & 0xFFFFFFFFull);
        slot = (uint32_t)ss.slot_of[v];
        ss.slot_of[v] = -1;
        ss.resident[v] = 0;
        ss.rowptr_tab[v] = nullptr;
        ss.col_tab[v] = nullptr;
        ds->victim = (int32_t)v;
    }
    uint8_t *base = ss.slots + (size_t)slot * ss.slot_bytes;
    const uint64_t n1 = (uint64_t)pt.node_off[c + 1] - pt.node_off[c] + 1;
    const uint64_t m = pt.edge_off[c + 1] - pt.edge_off[c];
    ss.slot_of[c] = (int32_t)slot;
    ss.part_in_slot[slot] = (int32_t)c;
    ss.resident[c] = 1;
    ss.rowptr_tab[c] = (const uint64_t *)base;
    ss.col_tab[c] = (const uint32_t *)(base + n1 * 8);
    ds->slot = (int32_t)slot;
    ds->load = 1;
    ds->loads++;
    ds->h2d_bytes += n1 * 8 * ss.rp_host + m * 4;
}

// Copy the decided partition from pinned host memory into its slot: rowptr
// slice, then col slice, as 4-byte words (grid-stride reads over PCIe).
__global__ __launch_bounds__(256) void load_kernel(SchedState ss, Parts pt) {
    const DevSched *ds = ss.ds;
    if (!ds->load) return;
    const uint32_t c = (uint32_t)ds->choice;
    const uint64_t n1 = (uint64_t)pt.node_off[c + 1] - pt.node_off[c] + 1;
    const uint64_t m = pt.edge_off[c + 1] - pt.edge_off[c];
    uint32_t *dst = (uint32_t *)(ss.slots + (size_t)ds->slot * ss.slot_bytes);
    const uint32_t *src_r = (const uint32_t *)(ss.zc_rowptr + pt.node_off[c]);
    const uint32_t *src_c = ss.zc_col + pt.edge_off[c];
    const uint64_t nr = 2 * n1, total = nr + m;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride)
        dst[i] = i < nr ? src_r[i] : src_c[i - nr];
}

// Unpacks a packed partition image into the col slice of an HBM slot whose
// rowptr slice is already in place (it gives the row boundaries; rowptr[0] is
// the partition's first global edge).  Image layout: meta u32[2*nch] (nch =
// ceil(m/64); per chunk its first word and the codes of its four 16-edge
// blocks), bit streams u32[nwords] + 2 zero words; an anchor (first edge of
// its row or of its chunk, or any edge of a block coded 255) holds the column
// in vbits bits, any other edge the gap to the previous column in its block's
// width.  The bytes written equal the raw col slice.
//
// One wave per kUnpackChunks consecutive 64-edge chunks, one lane per edge:
//   - the row holding the wave's first edge comes from a 64-ary search of
//     rowptr (one ballot per level); later chunks carry it forward;
//   - whether a lane's edge starts a row: rows ending inside the chunk are
//     read 64 at a time (coalesced) and mark their end in the wave's LDS
//     flags;
//   - each lane's bit offset: exclusive scan of the field widths; the chunk's
//     words (at most 64) sit one per lane and are fetched by shuffle;
//   - columns: inclusive scan of the fields minus the scan just before the
//     lane's anchor (u32 arithmetic; the difference is exact);
//   - the 64 columns of a chunk are stored by consecutive lanes.
constexpr uint32_t kUnpackChunks = 8;

// Inclusive prefix sum over the 64 lanes of a wave with DPP moves: shifts by
// 1, 2, 4 and 8 within each row of 16 lanes (lanes shifted in from outside
// the row read 0), then row 0's total is added to row 1 and row 2's to row 3
// (row_bcast:15), then lane 31's sum to rows 2 and 3 (row_bcast:31).
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x, uint32_t) {
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, true);  // row_shr:1
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, true);  // row_shr:2
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, true);  // row_shr:4
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, true);  // row_shr:8
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false); // row_bcast:15, rows 1, 3
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false); // row_bcast:31, rows 2, 3
    return x;
}

__global__ __launch_bounds__(256) void unpack_kernel(const uint8_t *__restrict__ img, uint32_t n, uint64_t m,
                                                     uint64_t nwords, uint32_t vbits,
                                                     const uint64_t *__restrict__ rowptr,
                                                     uint32_t *__restrict__ col) {
    const uint64_t nch = (m + 63) / 64;
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t cb = wave * kUnpackChunks;
    if (cb >= nch) return;  // whole wave
    const uint64_t ce = cb + kUnpackChunks < nch ? cb + kUnpackChunks : nch;
    const uint64_t eb = rowptr[0];
    const uint32_t *meta = (const uint32_t *)img;
    const uint32_t *wd = meta + 2 * nch;
    __shared__ uint8_t s_starts[4][64];  // per wave: chunk edges that start a row
    volatile uint8_t *starts = s_starts[threadIdx.x >> 6];

    // row r0 holding edge 64*cb: the last r < n with rowptr[r] - eb <= 64*cb
    uint64_t e0 = cb * 64;
    uint32_t lo = 0, hi = n - 1;
    while (hi > lo) {
        const uint64_t span = (uint64_t)hi - lo + 1;
        const uint32_t idx = lo + (uint32_t)(span * lane / 64);
        const uint64_t mask = __ballot(rowptr[idx] - eb <= e0);
        const uint32_t k = (uint32_t)__popcll(mask) - 1;  // lane 0 (idx = lo) always holds
        const uint32_t nlo = __shfl(idx, k, 64);
        const uint32_t nxt = __shfl(idx, k < 63 ? k + 1 : 63, 64);
        hi = k < 63 ? nxt - 1 : hi;
        lo = nlo;
        if (hi < lo) hi = lo;
    }
    uint32_t r0 = lo;

    for (uint64_t c = cb; c < ce; c++) {
        e0 = c * 64;
        const uint32_t cnt = (uint32_t)(m - e0 < 64 ? m - e0 : 64);
        // the chunk's words, one per lane, and its block codes
        const uint64_t w_beg = meta[2 * c], w_end = c + 1 < nch ? meta[2 * c + 2] : nwords;
        const uint32_t bw = (meta[2 * c + 1] >> ((lane >> 4) * 8)) & 0xFFu;
        const uint32_t word = w_beg + lane < w_end ? wd[w_beg + lane] : 0u;
        // rows r0+1.. that end inside the chunk: a row ending at chunk edge i
        // means edge i starts a row; mark those edges in the wave's LDS flags
        uint32_t adv = 0;
        starts[lane] = 0;
        for (uint32_t base = r0 + 1;; base += 64) {
            const uint32_t idx = base + lane;
            const uint64_t end = idx <= n ? rowptr[idx] - eb : ~0ull;  // > e0 for every idx > r0
            const uint32_t p = (uint32_t)(end - e0 < 65 ? end - e0 : 65);  // 1..65
            if (p < 64) starts[p] = 1;
            const uint32_t p63 = (uint32_t)__builtin_amdgcn_readlane((int)p, 63);
            adv += (uint32_t)__popcll(__ballot(p <= 64));
            if (p63 > 64 || base + 64 > n) break;
        }
        const bool valid = lane < cnt;
        const bool anchor = valid && (lane == 0 || starts[lane] || bw == 0xFFu);
        const uint32_t len = valid ? (anchor ? vbits : bw) : 0u;
        const uint32_t pos = wave_incl_scan(len, lane) - len;
        const uint32_t wi = pos >> 5;
        const uint32_t lo_w = __shfl(word, wi, 64);
        const uint32_t hi_w = __shfl(word, (wi + 1) & 63u, 64);
        // bits [pos, pos + len) of the chunk: (hi:lo) >> (pos % 32), low len bits
        const uint32_t bits = __builtin_amdgcn_alignbit(wi + 1 < 64 ? hi_w : 0u, lo_w, pos & 31u);
        const uint32_t v = len >= 32 ? bits : bits & ((1u << len) - 1u);
        const uint32_t incl = wave_incl_scan(v, lane);
        const uint64_t heads = __ballot(anchor) & (lane == 63 ? ~0ull : ((2ull << lane) - 1));
        const uint32_t head = 63u - (uint32_t)__clzll(heads);  // lane 0 is always a head
        const uint32_t before = __shfl(incl, head > 0 ? head - 1 : 0, 64);
        if (valid) col[e0 + lane] = incl - (head > 0 ? before : 0u);
        r0 += adv;
    }
}

}  // namespace lt
