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

// LDS layout of the append stage.  Every destination of a pass gets a slot:
// in direct mode the slot is the partition id (P slots); in hashed mode, for
// partition tables too large for that, the distinct destinations of the pass
// are numbered through an LDS hash table, so the layout depends only on the
// walkers per pass (at most one slot each).
constexpr uint32_t kDirectMaxP = 1024;  // larger partition tables use hashed slots
struct AppendLds {
    uint32_t *hist;   // [nslot] walkers per slot, then the slot's first reserved queue slot
    uint32_t *scan;   // [nslot+1] exclusive prefix of the walkers per slot
    uint32_t *ch0;    // [nslot] chunk holding the first reserved queue slot
    uint32_t *ch1;    // [nslot] chunk holding the last reserved queue slot
    uint32_t *part;   // hashed: [nslot] partition of each slot in use
    uint32_t *keys;   // hashed: [2*nslot] hash table of destination partitions (kEmpty = free)
    uint32_t *tslot;  // hashed: [2*nslot] slot given to each table entry
    uint32_t *nused;  // hashed: slots in use in the pass
    Walker *stage;    // [kBlock * wpt] walkers grouped by slot
    uint16_t *dst;    // [kBlock * wpt] slot of each staged walker
    uint32_t nslot;
    bool hashed;
};

// walkers staged per workgroup pass = kBlock * wpt
__host__ __device__ inline size_t append_lds_bytes(uint32_t P, int wpt) {
    const size_t nw = (size_t)kBlock * wpt;
    size_t b = P <= kDirectMaxP ? (4 * (size_t)P + 1) * 4 : (9 * nw + 2) * 4;
    b = (b + 15) & ~(size_t)15;
    return b + nw * (16 + 2);
}

__device__ __forceinline__ AppendLds carve_lds(uint8_t *smem, uint32_t P, int wpt) {
    AppendLds L;
    const uint32_t nw = kBlock * (uint32_t)wpt;
    L.hashed = P > kDirectMaxP;
    L.nslot = L.hashed ? nw : P;
    L.hist = (uint32_t *)smem;
    L.scan = L.hist + L.nslot;
    L.ch0 = L.scan + L.nslot + 1;
    L.ch1 = L.ch0 + L.nslot;
    L.part = L.ch1 + L.nslot;      // hashed mode only from here on
    L.keys = L.part + L.nslot;
    L.tslot = L.keys + 2 * L.nslot;
    L.nused = L.tslot + 2 * L.nslot;
    const size_t words = L.hashed ? 9 * (size_t)nw + 2 : 4 * (size_t)P + 1;
    const size_t off = (words * 4 + 15) & ~(size_t)15;
    L.stage = (Walker *)(smem + off);
    L.dst = (uint16_t *)(L.stage + nw);
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

// Store walker w at position rank of slot sl's reservation (its first queue
// slot in L.hist[sl], the chunks it touches in L.ch0/L.ch1).
__device__ __forceinline__ void store_reserved(const Queues &q, const AppendLds &L, uint32_t sl, uint32_t rank,
                                               const Walker &w) {
    const uint32_t b = L.hist[sl];
    const uint32_t s = b + rank;
    const uint32_t c = (s >> kChunkLog2) == (b >> kChunkLog2) ? L.ch0[sl] : L.ch1[sl];
    if (c < q.nchunks) q.slab[(size_t)c * kChunk + (s & (kChunk - 1))] = w;
    else atomicOr(&q.ctl->overflow, 8u);
}

// One queue reservation per walker (append_mode 2).  Collective like
// block_append: every walker first takes its queue slot; the walkers whose
// slot starts a chunk link it, and after the barrier the others wait
// (bounded) for links made by reservations of other workgroups.
template <int WPT>
__device__ void append_one_level(const Queues &q, const Walker (&w)[WPT], const uint32_t (&dst)[WPT]) {
    uint32_t b[WPT], c[WPT];
#pragma unroll
    for (int k = 0; k < WPT; k++) {
        c[k] = kEmpty;
        if (dst[k] == kEmpty) continue;
        b[k] = atomicAdd(&qlen_of(q, dst[k]), 1u);
        if ((b[k] & (kChunk - 1)) == 0) c[k] = alloc_chunk(q, dst[k], b[k] >> kChunkLog2);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < WPT; k++) {
        if (dst[k] == kEmpty) continue;
        if ((b[k] & (kChunk - 1)) != 0) c[k] = get_chunk(q, dst[k], b[k] >> kChunkLog2);
        if (c[k] < q.nchunks) q.slab[(size_t)c[k] * kChunk + (b[k] & (kChunk - 1))] = w[k];
        else atomicOr(&q.ctl->overflow, 8u);
    }
    __syncthreads();
}

// block_append for partition tables above kDirectMaxP: the pass's distinct
// destinations are numbered through an open-addressed LDS table of
// 2 * nslot entries (at most nslot keys, so every probe sequence finds its
// key or a free entry), and the histogram, scan, reservations and staging of
// block_append then work on those slots instead of on partition ids.
template <int WPT>
__device__ void block_append_hashed(const Queues &q, const AppendLds &L, const Walker (&w)[WPT],
                                    const uint32_t (&dst)[WPT], uint32_t *s_wsum, int mode) {
    const uint32_t tid = threadIdx.x;
    const uint32_t T = 2 * L.nslot;  // a power of two
    for (uint32_t i = tid; i < T; i += kBlock) L.keys[i] = kEmpty;
    if (tid == 0) *L.nused = 0;
    __syncthreads();
    uint32_t ent[WPT] = {};
#pragma unroll
    for (int k = 0; k < WPT; k++) {
        if (dst[k] == kEmpty) continue;
        uint32_t h = dst[k] * 2654435761u;
        h = (h ^ (h >> 15)) & (T - 1);
        for (uint32_t probe = 0; probe < T; probe++) {
            const uint32_t old = atomicCAS(&L.keys[h], kEmpty, dst[k]);
            if (old == kEmpty || old == dst[k]) break;
            h = (h + 1) & (T - 1);
        }
        ent[k] = h;
    }
    __syncthreads();
    for (uint32_t i = tid; i < T; i += kBlock) {
        const uint32_t key = L.keys[i];
        if (key != kEmpty) {
            const uint32_t sidx = atomicAdd(L.nused, 1u);
            L.tslot[i] = sidx;
            L.part[sidx] = key;
            L.hist[sidx] = 0;
        }
    }
    __syncthreads();
    const uint32_t n = *L.nused;
    uint32_t sl[WPT], rank[WPT];
#pragma unroll
    for (int k = 0; k < WPT; k++) {
        sl[k] = dst[k] != kEmpty ? L.tslot[ent[k]] : kEmpty;
        if (dst[k] != kEmpty) rank[k] = atomicAdd(&L.hist[sl[k]], 1u);
    }
    __syncthreads();
    const uint32_t total = block_exclusive_scan(L.hist, L.scan, n, s_wsum);
    if (tid == 0) L.scan[n] = total;
    __syncthreads();
    for (uint32_t i = tid; i < n; i += kBlock) {  // pass 1: the chunks this workgroup owns
        const uint32_t c = L.hist[i];
        if (c) {
            const uint32_t p = L.part[i];
            const uint32_t b = atomicAdd(&qlen_of(q, p), c);
            const uint32_t k0 = b >> kChunkLog2, k1 = (b + c - 1) >> kChunkLog2;
            const uint32_t c0 = (b & (kChunk - 1)) == 0 ? alloc_chunk(q, p, k0) : kEmpty;
            L.ch0[i] = c0;
            L.ch1[i] = k1 != k0 ? alloc_chunk(q, p, k1) : c0;
            L.hist[i] = b;
        }
    }
    for (uint32_t i = tid; i < n; i += kBlock) {  // pass 2: chunks owned by earlier reservations
        const uint32_t b = L.hist[i];
        if (L.scan[i + 1] != L.scan[i] && (b & (kChunk - 1)) != 0) {
            const uint32_t c0 = get_chunk(q, L.part[i], b >> kChunkLog2);
            const uint32_t k1 = (b + (L.scan[i + 1] - L.scan[i]) - 1) >> kChunkLog2;
            L.ch0[i] = c0;
            if (k1 == (b >> kChunkLog2)) L.ch1[i] = c0;
        }
    }
    __syncthreads();
    if (mode == 1) {  // uniform: each walker stored from its own lane
#pragma unroll
        for (int k = 0; k < WPT; k++)
            if (dst[k] != kEmpty) store_reserved(q, L, sl[k], rank[k], w[k]);
        __syncthreads();
        return;
    }
#pragma unroll
    for (int k = 0; k < WPT; k++) {
        if (dst[k] != kEmpty) {
            const uint32_t pos = L.scan[sl[k]] + rank[k];
            L.stage[pos] = w[k];
            L.dst[pos] = (uint16_t)sl[k];
        }
    }
    __syncthreads();
    for (uint32_t i = tid; i < total; i += kBlock) {
        const uint32_t slot = L.dst[i];
        const uint32_t b = L.hist[slot];
        const uint32_t s = b + (i - L.scan[slot]);
        const uint32_t c = (s >> kChunkLog2) == (b >> kChunkLog2) ? L.ch0[slot] : L.ch1[slot];
        if (c < q.nchunks) q.slab[(size_t)c * kChunk + (s & (kChunk - 1))] = L.stage[i];
        else atomicOr(&q.ctl->overflow, 8u);
    }
    __syncthreads();
}

// Append the walkers with dst[k] != kEmpty to their partitions' queues.
// Collective: every thread of the workgroup must call it.  Chunk ids are
// resolved once per (workgroup, destination): a reservation of at most
// kBlockWalkers <= kChunk slots touches at most two chunks; the range that
// contains a chunk's first slot links it, and a later range that arrives
// before that link is published links it itself (the CAS loser's chunk goes
// to the lost list).  Above kDirectMaxP partitions the slots are numbered
// through a hash table (block_append_hashed).
template <int WPT>
__device__ void block_append(const Queues &q, uint32_t P, const AppendLds &L, const Walker (&w)[WPT],
                             const uint32_t (&dst)[WPT], uint32_t *s_wsum, int mode = 0) {
    if (mode == 2) {  // uniform
        append_one_level<WPT>(q, w, dst);
        return;
    }
    if (L.hashed) {  // uniform: more partitions than LDS slots
        block_append_hashed<WPT>(q, L, w, dst, s_wsum, mode);
        return;
    }
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
    if (mode == 1) {  // uniform: each walker stored from its own lane
#pragma unroll
        for (int k = 0; k < WPT; k++)
            if (dst[k] != kEmpty) store_reserved(q, L, dst[k], rank[k], w[k]);
        __syncthreads();
        return;
    }
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
    uint32_t passes = 0;  // workgroup passes since the launch began (uniform in a workgroup)
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

// Column of partition-local edge e.  A packed slot holds, per block of 16
// edges, {base, bit offset} (the next block's offset gives the width) and
// the bit streams: column = base + bits(offset + (e % 16) * width, width).
__device__ __forceinline__ uint32_t span_col(const Span &sp, uint64_t e) {
    if (!sp.words) return sp.col[e];
    const uint64_t k = e >> 4;
    const uint2 h = ((const uint2 *)sp.col)[k];
    const uint32_t w = (sp.col[2 * k + 3] - h.y) >> 4;
    const uint32_t pos = h.y + (uint32_t)(e & 15) * w;
    const uint32_t lo = sp.words[pos >> 5], hi = sp.words[(pos >> 5) + 1];
    const uint32_t bits = __builtin_amdgcn_alignbit(hi, lo, pos & 31u);
    return h.x + (w >= 32 ? bits : bits & ((1u << w) - 1u));
}

__device__ __forceinline__ uint64_t walker_id(const Walker &w) { return ((uint64_t)w.id_hi << 32) | w.id_lo; }

// Visits go to a workgroup table in LDS (open addressing, 4 probes) and reach
// the global counters once per vertex every kVisitFlushPasses passes and at
// the end of the launch (flush_visits); a vertex that finds no slot is
// counted in global memory directly.  Flushing every pass instead cost 2 % of
// the walk time of R-MAT s26 pagerank with an 8 GB cache; every 16, 64 passes
// or only at the end measured the same.
constexpr uint32_t kVisitSlots = 2048;
constexpr uint32_t kVisitFlushPasses = 16;  // a power of two
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

// One update of app/pagerank.cu:26-44: count the visit, then teleport with
// probability 0.15 or move to a uniform neighbour.  A vertex without
// neighbours teleports (the reference divides by zero there).
__device__ __forceinline__ void step_pagerank(Walker &w, const Span &sp, const AppArgs &a, Counters &ct) {
    count_visit(a, w.cur, ct);
    ct.visits++;
    w.step -= 1;
    if (w.step == 0) return;
    const u32x4 x = draw(a.seed, walker_id(w), w.step, TAG_PAGERANK);
    bool tele = x.x < a.thresh;
    if (!tele) {
        const uint64_t *r = sp.rowptr + (w.cur - sp.nb);
        const uint64_t b = r[0], d = r[1] - b;
        if (d == 0) {
            tele = true;
            if (sp.zc) ct.zc_reads += a.rp_host;
        } else {
            w.cur = span_col(sp, b - sp.eb + bounded(x.y, x.z, d));
            ct.moves++;
            if (sp.zc) { ct.zc_moves++; ct.zc_reads += 1 + a.rp_host; }
        }
    }
    if (tele) {
        w.cur = (uint32_t)bounded(x.y, x.z, a.V);
        ct.tele++;
    }
}

// One update of app/ppr.cu:22-42: the endpoint is counted when one step is
// left; a vertex without neighbours is counted and ends the walk.
__device__ __forceinline__ void step_ppr(Walker &w, const Span &sp, const AppArgs &a, Counters &ct) {
    if (w.step == 1) {
        if (a.counts) count_visit(a, w.cur, ct);
        ct.visits++;
    }
    w.step -= 1;
    if (w.step == 0) return;
    const uint64_t *r = sp.rowptr + (w.cur - sp.nb);
    const uint64_t b = r[0], d = r[1] - b;
    if (d > 0) {
        const u32x4 x = draw(a.seed, walker_id(w), w.step, TAG_PPR);
        w.cur = span_col(sp, b - sp.eb + bounded(x.y, x.z, d));
        ct.moves++;
        if (sp.zc) { ct.zc_moves++; ct.zc_reads += 1 + a.rp_host; }
    } else {
        if (a.counts) count_visit(a, w.cur, ct);
        ct.visits++;
        w.step = 0;
        if (sp.zc) ct.zc_reads += a.rp_host;
    }
}

// One update of app/genericwalk.cu:22-32: a uniform neighbour move, or stop
// at a vertex without neighbours.  Records the number of moves made.
__device__ __forceinline__ void step_generic(Walker &w, const Span &sp, const AppArgs &a, Counters &ct) {
    const uint64_t *r = sp.rowptr + (w.cur - sp.nb);
    const uint64_t b = r[0], d = r[1] - b;
    uint32_t *path = a.paths ? a.paths + walker_id(w) * ((uint64_t)a.length + 1) : nullptr;
    if (path) path[a.length - w.step] = w.cur;  // the vertex after (length - step) moves
    if (d > 0) {
        const u32x4 x = draw(a.seed, walker_id(w), w.step, TAG_GENERIC);
        w.cur = span_col(sp, b - sp.eb + bounded(x.y, x.z, d));
        ct.moves++;
        if (sp.zc) { ct.zc_moves++; ct.zc_reads += 1 + a.rp_host; }
        w.step -= 1;
        if (w.step == 0) {
            if (a.moves_out) a.moves_out[walker_id(w)] = a.length;
            if (path) path[a.length] = w.cur;
        }
    } else {
        if (a.moves_out) a.moves_out[walker_id(w)] = a.length - w.step;
        w.step = 0;
        if (sp.zc) ct.zc_reads += a.rp_host;
    }
}

template <int APP>
__device__ __forceinline__ void app_step(Walker &w, const Span &sp, const AppArgs &a, Counters &ct) {
    if (APP == APP_PAGERANK) step_pagerank(w, sp, a, ct);
    else if (APP == APP_PPR) step_ppr(w, sp, a, ct);
    else step_generic(w, sp, a, ct);
}

__device__ __forceinline__ void finish_walker(const Walker &w, const AppArgs &a) {
    if (a.endpoints) a.endpoints[walker_id(w)] = w.cur;
}

// Walk one residency; returns the destination partition or kEmpty when the
// walk is over.  With span_resident the walker continues into any partition
// whose slot is resident (and into the source slice itself).
template <int APP>
__device__ uint32_t walk_residency(Walker &w, const WalkArgs &a, uint32_t src, const Span &src_sp, Counters &ct) {
    Span sp = src_sp;
    while (true) {
        while (w.step > 0 && (w.cur - sp.nb) < (sp.ne - sp.nb)) app_step<APP>(w, sp, a.app, ct);
        if (w.step == 0) {
            finish_walker(w, a.app);
            return kEmpty;
        }
        const uint32_t p = find_part(a.pt, w.cur);
        if (!a.span_resident) return p;
        if (p == src) {
            sp = src_sp;
            continue;
        }
        const uint64_t *rp = a.pt.rowptr_tab[p];
        if (!rp) return p;
        sp.rowptr = rp;
        sp.col = a.pt.col_tab[p];
        sp.words = a.pt.words_tab[p];
        sp.nb = a.pt.node_off[p];
        sp.ne = a.pt.node_off[p + 1];
        sp.eb = a.pt.edge_off[p];
        sp.zc = a.pt.zc_part[p];
    }
}

// One workgroup pass over nvalid <= kBlock * WPT walkers stored
// contiguously at cw: walk each, then append the survivors to their queues.
template <int APP, int WPT>
__device__ __forceinline__ void walk_walkers(const WalkArgs &a, const AppendLds &L, uint32_t *s_wsum, uint32_t src,
                                            const Span &src_sp, const Walker *cw, uint32_t nvalid, Counters &ct) {
    Walker w[WPT];
    uint32_t dst[WPT];
#pragma unroll
    for (int k = 0; k < WPT; k++) {
        const uint32_t i = (uint32_t)k * kBlock + threadIdx.x;
        dst[k] = kEmpty;
        if (i < nvalid) {
            w[k] = cw[i];
            ct.resid++;
        } else {
            w[k].step = 0;
        }
    }
#pragma unroll
    for (int k = 0; k < WPT; k++) {
        if ((uint32_t)k * kBlock + threadIdx.x < nvalid) dst[k] = walk_residency<APP>(w[k], a, src, src_sp, ct);
    }
    block_append<WPT>(a.q, a.pt.P, L, w, dst, s_wsum, a.append_mode);
    if constexpr (APP != APP_GENERIC) {  // genericwalk counts no visits
        if ((++ct.passes & (kVisitFlushPasses - 1)) == 0) flush_visits(a.app, ct);
    }
}

// One workgroup pass: walkers [base, base + kBlock * WPT) of src's queue,
// clipped to hi64; base is a multiple of kBlock * WPT, so the pass reads one
// chunk.
template <int APP, int WPT>
__device__ __forceinline__ void walk_pass(const WalkArgs &a, const AppendLds &L, uint32_t *s_wsum, uint32_t src,
                                          const Span &src_sp, uint64_t base, uint64_t hi64, Counters &ct) {
    const uint32_t chunk = a.q.frame[(size_t)src * a.q.maxc + (uint32_t)(base >> kChunkLog2)];
    const bool chunk_ok = chunk < a.q.nchunks;
    if (!chunk_ok && threadIdx.x == 0) atomicOr(&a.q.ctl->overflow, 8u);
    const Walker *cw = a.q.slab + (size_t)(chunk_ok ? chunk : 0) * kChunk + (uint32_t)(base & (kChunk - 1));
    const uint64_t left = hi64 > base ? hi64 - base : 0;
    const uint32_t nvalid = chunk_ok ? (uint32_t)min(left, (uint64_t)(kBlock * WPT)) : 0u;
    walk_walkers<APP, WPT>(a, L, s_wsum, src, src_sp, cw, nvalid, ct);
}

// WPT walkers per thread per pass (1, 2 or 4): small queues use 1 so that
// enough workgroups exist to fill the chip, large ones 4.
template <int APP, int WPT>
__global__ __launch_bounds__(kBlock) void walk_kernel(WalkArgs a) {
    constexpr int BW = kBlock * WPT;
    extern __shared__ __attribute__((aligned(16))) uint8_t smem[];
    __shared__ uint32_t s_wsum[kBlock / 64];
    const AppendLds L = carve_lds(smem, a.pt.P, WPT);
    Counters ct;
    if constexpr (APP != APP_GENERIC) {
        __shared__ uint32_t s_vkeys[kVisitSlots], s_vcnt[kVisitSlots];
        if (a.app.counts) {  // uniform
            for (uint32_t i = threadIdx.x; i < kVisitSlots; i += blockDim.x) {
                s_vkeys[i] = kVisitEmpty;
                s_vcnt[i] = 0;
            }
            __syncthreads();
            ct.vkeys = s_vkeys;
            ct.vcnt = s_vcnt;
        }
    }
    if (a.nsrc) {  // batch: the whole queues of several partitions, each in the partition table
        const uint64_t vend = a.voff[a.nsrc];
        for (uint64_t base = (uint64_t)blockIdx.x * BW; base < vend; base += (uint64_t)gridDim.x * BW) {
            uint32_t k = 0;
            while (k + 1 < a.nsrc && base >= a.voff[k + 1]) k++;
            const uint32_t src = a.srcs[k];
            Span sp;  // as a walker entering src reads it (walk_residency)
            sp.rowptr = a.pt.rowptr_tab[src];
            sp.col = a.pt.col_tab[src];
            sp.words = a.pt.words_tab[src];
            sp.nb = a.pt.node_off[src];
            sp.ne = a.pt.node_off[src + 1];
            sp.eb = a.pt.edge_off[src];
            sp.zc = a.pt.zc_part[src];
            const uint64_t local = base - a.voff[k], n = qlen_of(a.q, src);
            if (local < n) walk_pass<APP, WPT>(a, L, s_wsum, src, sp, local, n, ct);  // uniform per workgroup
        }
        if constexpr (APP != APP_GENERIC) flush_visits(a.app, ct);
        flush_counters(ct, a.stats);
        return;
    }
    uint32_t src = a.src, chunk_lo = a.chunk_lo, chunk_hi = a.chunk_hi;
    Span src_sp;
    src_sp.rowptr = a.rowptr;
    src_sp.col = a.col;
    src_sp.words = a.words;
    src_sp.zc = a.zerocopy;
    if (a.nhost) {  // walkers of src spilled to host memory, read over PCIe
        src_sp.nb = a.pt.node_off[src];
        src_sp.ne = a.pt.node_off[src + 1];
        src_sp.eb = a.pt.edge_off[src];
        const uint64_t vend = (uint64_t)a.nhost << kChunkLog2;
        for (uint64_t base = (uint64_t)blockIdx.x * BW; base < vend; base += (uint64_t)gridDim.x * BW) {
            const uint32_t k = (uint32_t)(base >> kChunkLog2), local = (uint32_t)(base & (kChunk - 1));
            if (local < a.hn[k])  // uniform per workgroup
                walk_walkers<APP, WPT>(a, L, s_wsum, src, src_sp, a.hptr[k] + local,
                                       min(a.hn[k] - local, (uint32_t)BW), ct);
        }
        if constexpr (APP != APP_GENERIC) flush_visits(a.app, ct);
        flush_counters(ct, a.stats);
        return;
    }
    if (a.ss.ds) {  // the decision of this iteration was taken on the device
        const int32_t c = a.ss.ds->choice;
        if (c < 0) return;
        src = (uint32_t)c;
        chunk_lo = 0;
        chunk_hi = 0xFFFFFFFFu;
        src_sp.zc = a.ss.ds->zerocopy;
        src_sp.words = nullptr;
        if (src_sp.zc) {
            src_sp.rowptr = a.ss.zc_rowptr + a.pt.node_off[src];
            src_sp.col = a.ss.zc_col + a.pt.edge_off[src];
        } else {  // a raw slot, or a block slot whose bit streams decide_kernel entered in the table
            const uint8_t *base = a.ss.slots + (size_t)a.ss.slot_of[src] * a.ss.slot_bytes;
            src_sp.rowptr = (const uint64_t *)base;
            src_sp.col = (const uint32_t *)(base + ((uint64_t)a.pt.node_off[src + 1] - a.pt.node_off[src] + 1) * 8);
            src_sp.words = a.pt.words_tab[src];
        }
    }
    const uint32_t n = qlen_of(a.q, src);
    const uint64_t lo64 = (uint64_t)chunk_lo << kChunkLog2;
    const uint64_t hi64 = min((uint64_t)n, (uint64_t)chunk_hi << kChunkLog2);
    src_sp.nb = a.pt.node_off[src];
    src_sp.ne = a.pt.node_off[src + 1];
    src_sp.eb = a.pt.edge_off[src];
    for (uint64_t base = lo64 + (uint64_t)blockIdx.x * BW; base < hi64; base += (uint64_t)gridDim.x * BW)
        walk_pass<APP, WPT>(a, L, s_wsum, src, src_sp, base, hi64, ct);
    if constexpr (APP != APP_GENERIC) flush_visits(a.app, ct);
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
__global__ __launch_bounds__(1024) void decide_kernel(Queues q, SchedState ss, Parts pt, BlockImages bi) {
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
        bi.words_tab[v] = nullptr;
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
    const uint64_t img = bi.img ? bi.size[c] : 0;  // block image bytes, 0 = raw columns
    bi.words_tab[c] = img ? ss.col_tab[c] + 2 * ((m + kBlockEdges - 1) / kBlockEdges + 1) : nullptr;
    ds->slot = (int32_t)slot;
    ds->load = 1;
    ds->loads++;
    ds->h2d_bytes += n1 * 8 * ss.rp_host + (img ? img : m * 4);
}

// Copy the decided partition from pinned host memory into its slot: rowptr
// slice, then the col slice or (block slots) the partition's block image, as
// 4-byte words (grid-stride reads over PCIe).
__global__ __launch_bounds__(256) void load_kernel(SchedState ss, Parts pt, BlockImages bi) {
    const DevSched *ds = ss.ds;
    if (!ds->load) return;
    const uint32_t c = (uint32_t)ds->choice;
    const uint64_t n1 = (uint64_t)pt.node_off[c + 1] - pt.node_off[c] + 1;
    const uint64_t img = bi.img ? bi.size[c] : 0;
    const uint64_t m = img ? img / 4 : pt.edge_off[c + 1] - pt.edge_off[c];  // words after the rowptr slice
    uint32_t *dst = (uint32_t *)(ss.slots + (size_t)ds->slot * ss.slot_bytes);
    const uint32_t *src_r = (const uint32_t *)(ss.zc_rowptr + pt.node_off[c]);
    const uint32_t *src_c = img ? (const uint32_t *)(bi.img + bi.off[c]) : ss.zc_col + pt.edge_off[c];
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
