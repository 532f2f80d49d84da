/*
 * lt_oracle.c — CPU oracle for the LightTraffic random-walk path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * engine (liblighttraffic.so) never links, loads or calls it.
 *
 * It restates, in plain C with OpenMP over walkers and with the whole graph
 * in DRAM (no partitions, no queues, no scheduler), the reference's walk
 * semantics with the engine's documented random draws:
 *   - pagerank step:    app/pagerank.cu:26-44
 *   - ppr step:         app/ppr.cu:22-42
 *   - genericwalk step: app/genericwalk.cu:22-32
 *   - uniform starts:   src/walk/NodeSampler.h:5-15
 *   - ppr source+len:   src/walk/SingleSource.h:6-20
 * and the host-side format/partition/scheduler rules:
 *   - edge list -> bcsr:     tools/convert_from_txt.cc:66-218
 *   - greedy partitioning:   src/partition/PartitionStrategy.h:189-235
 *   - findPartition:         src/partition/PartitionStrategy.h:151-168
 *   - scheduler decisions:   src/Scheduler.h:84-171, 368-388
 *
 * Random draws: Philox4x32-10 keyed by the seed, counter (id_lo, id_hi,
 * ctr, tag).  The reference uses curand (non-reproducible, see
 * src/Rand.h:28-30), so the oracle fixes the engine's own integer rules:
 * bounded(hi, lo, n) = high64(((hi<<32)|lo) * n), Bernoulli(p) = r < floor(p*2^32).
 * Philox is implemented here independently of the engine's lt_rng.h and is
 * pinned by the Random123 known-answer vectors in tests/golden.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define TAG_START 0x53545254u
#define TAG_PPRLEN 0x504C454Eu
#define TAG_PAGERANK 0x50524B53u
#define TAG_PPR 0x50505253u
#define TAG_GENERIC 0x47454E53u
#define TAG_RMAT 0x524D4154u

/* ---- Philox4x32-10 (Salmon et al., SC'11, Random123 philox.h) ---------- */
void ltor_philox(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int round = 0; round < 10; round++) {
        if (round > 0) {
            k0 = k0 + 0x9E3779B9u; /* Weyl key schedule */
            k1 = k1 + 0xBB67AE85u;
        }
        uint64_t prod0 = (uint64_t)c0 * 0xD2511F53u;
        uint64_t prod1 = (uint64_t)c2 * 0xCD9E8D57u;
        uint32_t hi0 = (uint32_t)(prod0 >> 32), lo0 = (uint32_t)prod0;
        uint32_t hi1 = (uint32_t)(prod1 >> 32), lo1 = (uint32_t)prod1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline void rnd(uint64_t seed, uint64_t id, uint32_t ctr, uint32_t tag, uint32_t o[4]) {
    uint32_t c[4] = {(uint32_t)(id & 0xFFFFFFFFu), (uint32_t)(id >> 32), ctr, tag};
    uint32_t k[2] = {(uint32_t)(seed & 0xFFFFFFFFu), (uint32_t)(seed >> 32)};
    ltor_philox(c, k, o);
}

/* uniform in [0, n): high 64 bits of the 128-bit product */
static inline uint64_t pick(uint32_t hi, uint32_t lo, uint64_t n) {
    unsigned __int128 r = ((unsigned __int128)(((uint64_t)hi << 32) | lo)) * n;
    return (uint64_t)(r >> 64);
}

uint32_t ltor_threshold(double p) {
    if (!(p > 0.0)) return 0u;
    if (p >= 1.0) return 0xFFFFFFFFu;
    double t = p * 4294967296.0;
    if (t >= 4294967295.0) return 0xFFFFFFFFu;
    return (uint32_t)t;
}

/* Reference pagerank uses alpha = 0.85f and teleports when
 * curand_uniform_double > alpha (app/pagerank.cu:6,35): probability
 * 1 - (double)0.85f. */
uint32_t ltor_teleport_threshold(void) { return ltor_threshold(1.0 - (double)0.85f); }

static inline uint64_t deg_of(const uint64_t *rowptr, uint32_t v) { return rowptr[v + 1] - rowptr[v]; }

/* stats[0]=moves, [1]=teleports, [2]=visits */
int ltor_pagerank(const uint64_t *rowptr, const uint32_t *col, uint32_t V, uint32_t l, uint64_t id_begin,
                  uint64_t id_end, uint64_t seed, uint32_t *counts, uint32_t *endpoints, int nthreads,
                  uint64_t *stats) {
    if (V == 0) return -1;
    const uint32_t T = ltor_teleport_threshold();
    uint64_t moves = 0, tele = 0, visits = 0;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : moves, tele, visits)
    for (int64_t k = (int64_t)id_begin; k < (int64_t)id_end; k++) {
        uint64_t id = (uint64_t)k;
        uint32_t r[4];
        rnd(seed, id, 0, TAG_START, r);
        uint32_t cur = (uint32_t)pick(r[0], r[1], V);
        uint32_t s = l;
        while (s > 0) {
            if (counts) {
#pragma omp atomic
                counts[cur]++;
            }
            visits++;
            s--;
            if (s == 0) break;
            rnd(seed, id, s, TAG_PAGERANK, r);
            uint64_t d = deg_of(rowptr, cur);
            if (r[0] < T || d == 0) {
                cur = (uint32_t)pick(r[1], r[2], V);
                tele++;
            } else {
                cur = col[rowptr[cur] + pick(r[1], r[2], d)];
                moves++;
            }
        }
        if (endpoints) endpoints[id - id_begin] = cur;
    }
    if (stats) { stats[0] = moves; stats[1] = tele; stats[2] = visits; }
    return 0;
}

/* Geometric length of a PPR walker (src/walk/SingleSource.h:12-17):
 * step = 2 + number of draws r >= T before the first r < T. */
static uint32_t ppr_length(uint64_t seed, uint64_t id, uint32_t T) {
    uint32_t F = 0;
    for (uint32_t j = 0;; j++) {
        uint32_t r[4];
        rnd(seed, id, j, TAG_PPRLEN, r);
        for (int q = 0; q < 4; q++) {
            if (r[q] < T) return 2u + F;
            F++;
        }
        if (F > 0xFFFFFF00u) return 2u + F;
    }
}

static uint32_t ppr_walk(const uint64_t *rowptr, const uint32_t *col, uint64_t seed, uint64_t id, uint32_t source,
                         uint32_t T, uint32_t *counts, uint64_t *moves) {
    uint32_t cur = source;
    uint32_t s = ppr_length(seed, id, T);
    uint64_t mv = 0;
    while (s > 0) {
        if (s == 1 && counts) {
#pragma omp atomic
            counts[cur]++;
        }
        s--;
        if (s == 0) break;
        uint64_t d = deg_of(rowptr, cur);
        if (d > 0) {
            uint32_t r[4];
            rnd(seed, id, s, TAG_PPR, r);
            cur = col[rowptr[cur] + pick(r[1], r[2], d)];
            mv++;
        } else {
            if (counts) {
#pragma omp atomic
                counts[cur]++;
            }
            s = 0;
        }
    }
    *moves += mv;
    return cur;
}

int ltor_ppr(const uint64_t *rowptr, const uint32_t *col, uint32_t V, uint32_t source, float prob, uint64_t id_begin,
             uint64_t id_end, uint64_t seed, uint32_t *counts, uint32_t *endpoints, int nthreads, uint64_t *stats) {
    if (source >= V) return -1;
    const uint32_t T = ltor_threshold((double)prob);
    if (T == 0) return -1;
    uint64_t moves = 0;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : moves)
    for (int64_t k = (int64_t)id_begin; k < (int64_t)id_end; k++) {
        uint32_t e = ppr_walk(rowptr, col, seed, (uint64_t)k, source, T, counts, &moves);
        if (endpoints) endpoints[(uint64_t)k - id_begin] = e;
    }
    if (stats) { stats[0] = moves; stats[1] = 0; stats[2] = id_end - id_begin; }
    return 0;
}

/* Batched PPR: query q owns walker ids [off_q, off_q + counts[q]). */
int ltor_ppr_batch(const uint64_t *rowptr, const uint32_t *col, uint32_t V, uint32_t nq, const uint32_t *sources,
                   const uint64_t *qcounts, float prob, uint64_t seed, uint32_t *endpoints, int nthreads,
                   uint64_t *stats) {
    const uint32_t T = ltor_threshold((double)prob);
    if (T == 0) return -1;
    uint64_t *off = (uint64_t *)malloc(sizeof(uint64_t) * (nq + 1));
    off[0] = 0;
    for (uint32_t q = 0; q < nq; q++) {
        if (sources[q] >= V) { free(off); return -1; }
        off[q + 1] = off[q] + qcounts[q];
    }
    uint64_t moves = 0;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : moves)
    for (int64_t q = 0; q < (int64_t)nq; q++) {
        for (uint64_t id = off[q]; id < off[q + 1]; id++)
            endpoints[id] = ppr_walk(rowptr, col, seed, id, sources[q], T, NULL, &moves);
    }
    if (stats) { stats[0] = moves; stats[1] = 0; stats[2] = off[nq]; }
    free(off);
    return 0;
}

int ltor_genericwalk(const uint64_t *rowptr, const uint32_t *col, uint32_t V, uint32_t l, uint64_t id_begin,
                     uint64_t id_end, uint64_t seed, uint32_t *endpoints, uint32_t *moves_out, int nthreads,
                     uint64_t *stats) {
    if (V == 0) return -1;
    uint64_t moves = 0;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : moves)
    for (int64_t k = (int64_t)id_begin; k < (int64_t)id_end; k++) {
        uint64_t id = (uint64_t)k;
        uint32_t r[4];
        rnd(seed, id, 0, TAG_START, r);
        uint32_t cur = (uint32_t)pick(r[0], r[1], V);
        uint32_t s = l, mv = 0;
        while (s > 0) {
            uint64_t d = deg_of(rowptr, cur);
            if (d > 0) {
                rnd(seed, id, s, TAG_GENERIC, r);
                cur = col[rowptr[cur] + pick(r[1], r[2], d)];
                mv++;
                s--;
            } else {
                s = 0;
            }
        }
        moves += mv;
        if (endpoints) endpoints[id - id_begin] = cur;
        if (moves_out) moves_out[id - id_begin] = mv;
    }
    if (stats) { stats[0] = moves; stats[1] = 0; stats[2] = 0; }
    return 0;
}

/* genericwalk paths: row id of paths (l + 1 entries) holds the vertex after
 * k moves at column k, the start at column 0, and 0xFFFFFFFF after a walk that
 * stopped early (the layout of sampleNodeWithId / pathSampler,
 * src/walk/NodeSampler.h:36-49, src/walk/PathSampler.h:6-19; the steps of
 * app/genericwalk.cu:22-32 as in ltor_genericwalk). */
int ltor_genericwalk_paths(const uint64_t *rowptr, const uint32_t *col, uint32_t V, uint32_t l, uint64_t id_begin,
                           uint64_t id_end, uint64_t seed, uint32_t *paths, int nthreads) {
    if (V == 0 || !paths) return -1;
    const uint64_t row = (uint64_t)l + 1;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t k = (int64_t)id_begin; k < (int64_t)id_end; k++) {
        uint64_t id = (uint64_t)k;
        uint32_t *p = paths + (id - id_begin) * row;
        for (uint64_t j = 0; j < row; j++) p[j] = 0xFFFFFFFFu;
        uint32_t r[4];
        rnd(seed, id, 0, TAG_START, r);
        uint32_t cur = (uint32_t)pick(r[0], r[1], V);
        uint32_t s = l, mv = 0;
        p[0] = cur;
        while (s > 0) {
            uint64_t d = deg_of(rowptr, cur);
            if (d == 0) break;
            rnd(seed, id, s, TAG_GENERIC, r);
            cur = col[rowptr[cur] + pick(r[1], r[2], d)];
            p[++mv] = cur;
            s--;
        }
    }
    return 0;
}

/* ---- R-MAT edge sampling ------------------------------------------------- */
/* Edge e, level t: word t%4 of Philox(seed; e, t/4, TAG_RMAT); quadrant by
 * cumulative thresholds of (a, a+b, a+b+c); level t sets bit (scale-1-t). */
int ltor_rmat_edges(uint32_t scale, uint64_t n_edges, double a, double b, double c, uint64_t seed, uint32_t *src,
                    uint32_t *dst, int nthreads) {
    if (scale == 0 || scale > 32) return -1;
    const uint32_t ta = ltor_threshold(a), tab = ltor_threshold(a + b), tabc = ltor_threshold(a + b + c);
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < (int64_t)n_edges; e++) {
        uint64_t s = 0, d = 0;
        uint32_t r[4];
        for (uint32_t t = 0; t < scale; t++) {
            if ((t & 3u) == 0) rnd(seed, (uint64_t)e, t >> 2, TAG_RMAT, r);
            uint32_t u = r[t & 3u];
            uint32_t sb, db;
            if (u < ta) { sb = 0; db = 0; }
            else if (u < tab) { sb = 0; db = 1; }
            else if (u < tabc) { sb = 1; db = 0; }
            else { sb = 1; db = 1; }
            s |= (uint64_t)sb << (scale - 1 - t);
            d |= (uint64_t)db << (scale - 1 - t);
        }
        src[e] = (uint32_t)s;
        dst[e] = (uint32_t)d;
    }
    return 0;
}

/* ---- edge list -> bcsr (tools/convert_from_txt.cc:66-218) --------------- */
static int cmp_u64(const void *x, const void *y) {
    uint64_t a = *(const uint64_t *)x, b = *(const uint64_t *)y;
    return a < b ? -1 : (a > b ? 1 : 0);
}

/* Drops self loops, relabels the surviving ids 0..max densely (isolated
 * vertices removed), adds reverse edges, sorts by (src,dst), removes
 * duplicates.  V = last source + 1.  Outputs are malloc'd; free with
 * ltor_free. */
int ltor_convert(uint64_t n, const uint64_t *src, const uint64_t *dst, uint32_t *V_out, uint64_t *E_out,
                 uint64_t **rowptr_out, uint32_t **col_out) {
    uint64_t max = 0, kept = 0;
    for (uint64_t i = 0; i < n; i++) {
        if (src[i] == dst[i]) continue;
        kept++;
        if (src[i] > max) max = src[i];
        if (dst[i] > max) max = dst[i];
    }
    uint8_t *existed = (uint8_t *)calloc(max + 1, 1);
    for (uint64_t i = 0; i < n; i++) {
        if (src[i] == dst[i]) continue;
        existed[src[i]] = 1;
        existed[dst[i]] = 1;
    }
    uint64_t *map = (uint64_t *)malloc(sizeof(uint64_t) * (max + 1));
    uint64_t cnt = 0;
    for (uint64_t v = 0; v <= max; v++) {
        map[v] = cnt;
        if (existed[v]) cnt++;
    }
    free(existed);
    uint64_t *keys = (uint64_t *)malloc(sizeof(uint64_t) * (2 * kept + 1));
    uint64_t m = 0;
    for (uint64_t i = 0; i < n; i++) {
        if (src[i] == dst[i]) continue;
        uint64_t s = map[src[i]], d = map[dst[i]];
        keys[m++] = (s << 32) | d;
        keys[m++] = (d << 32) | s;
    }
    free(map);
    qsort(keys, m, sizeof(uint64_t), cmp_u64);
    uint64_t u = 0;
    for (uint64_t i = 0; i < m; i++)
        if (u == 0 || keys[i] != keys[u - 1]) keys[u++] = keys[i];
    uint32_t V = u ? (uint32_t)(keys[u - 1] >> 32) + 1 : 0;
    uint64_t *rowptr = (uint64_t *)calloc((uint64_t)V + 1, sizeof(uint64_t));
    uint32_t *col = (uint32_t *)malloc(sizeof(uint32_t) * (u ? u : 1));
    for (uint64_t i = 0; i < u; i++) {
        rowptr[(keys[i] >> 32) + 1]++;
        col[i] = (uint32_t)(keys[i] & 0xFFFFFFFFu);
    }
    for (uint32_t v = 0; v < V; v++) rowptr[v + 1] += rowptr[v];
    free(keys);
    *V_out = V;
    *E_out = u;
    *rowptr_out = rowptr;
    *col_out = col;
    return 0;
}

void ltor_free(void *p) { free(p); }

/* ---- greedy partitioning (src/partition/PartitionStrategy.h:189-235) ---- */
/* Returns P >= 1, -1 when a single vertex exceeds max_bytes (with *bad set to
 * that vertex), -2 when cap is too small. */
int64_t ltor_partition(const uint64_t *rowptr, uint32_t V, uint64_t max_bytes, uint32_t *node_off,
                       uint64_t *edge_off, uint64_t cap, uint32_t *bad) {
    uint64_t P = 0;
    uint32_t n = 0;
    if (cap < 2) return -2;
    node_off[0] = 0;
    edge_off[0] = 0;
    for (uint64_t i = 1; i <= V;) {
        uint64_t bytes = (i - n + 1) * 8u + (rowptr[i] - rowptr[n]) * 4u;
        if (bytes > max_bytes) {
            if (i == (uint64_t)n + 1) {
                if (bad) *bad = n;
                return -1;
            }
            n = (uint32_t)(i - 1);
            P++;
            if (P + 1 >= cap) return -2;
            node_off[P] = n;
            edge_off[P] = rowptr[n];
        } else {
            i++;
        }
    }
    P++;
    node_off[P] = V;
    edge_off[P] = rowptr[V];
    return (int64_t)P;
}

/* findPartition (src/partition/PartitionStrategy.h:151-168). */
uint32_t ltor_find_partition(const uint32_t *node_off, uint32_t P, uint32_t v) {
    uint32_t low = 0, high = P - 1;
    while (low < high) {
        uint32_t mid = low + (high - low + 1) / 2;
        if (v < node_off[mid]) high = mid - 1;
        else low = mid;
    }
    return low;
}

/* ---- scheduler decisions ------------------------------------------------ */
/* Scheduler::nextPartition, greedy (src/Scheduler.h:368-388). */
int32_t ltor_greedy_next(uint32_t P, const uint64_t *counts, int32_t prev, int32_t sync) {
    uint64_t max = 0;
    int32_t choice = -1;
    for (uint32_t i = 0; i < P; i++) {
        if ((int32_t)i == prev && !sync) continue;
        if (max < counts[i]) { max = counts[i]; choice = (int32_t)i; }
    }
    return choice;
}

/* findPartitionwithleastWalker (src/Scheduler.h:149-171), ties -> lowest id. */
int32_t ltor_evict_min(uint32_t P, const uint64_t *counts, const uint8_t *resident, int32_t except) {
    int32_t choice = -1;
    uint64_t min = 0;
    for (uint32_t i = 0; i < P; i++) {
        if (!resident[i] || (int32_t)i == except) continue;
        if (choice < 0 || counts[i] < min) { min = counts[i]; choice = (int32_t)i; }
    }
    return choice;
}

/* getComputedPartition, greedy variant (src/Scheduler.h:84-111). */
int32_t ltor_computed_partition(uint32_t P, const uint64_t *counts, const uint8_t *resident, int32_t except,
                                uint64_t page) {
    int32_t full = -1, notfull = -1;
    uint64_t min_full = 0, max_notfull = 0;
    for (uint32_t i = 0; i < P; i++) {
        if ((int32_t)i == except || !resident[i]) continue;
        uint64_t n = counts[i];
        if (n >= page) {
            if (full < 0 || n < min_full) { full = (int32_t)i; min_full = n; }
        } else if (full < 0 && n > 0) {
            if (notfull < 0 || n > max_notfull) { notfull = (int32_t)i; max_notfull = n; }
        }
    }
    return full >= 0 ? full : notfull;
}
