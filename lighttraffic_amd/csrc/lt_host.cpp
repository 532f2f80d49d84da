// lt_host.cpp — graph and partition files, host partitioner, pure scheduler
// rules, status strings.  No kernels live here.
#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <math.h>

#include <algorithm>
#include <functional>
#include <memory>
#include <set>
#include <vector>

#include "lt_internal.h"
#include "lt_rng.h"

namespace lt {

static thread_local char g_err[512];

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

static bool device_present() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return n > 0;
}

Graph::~Graph() {
    if (pinned) {
        if (rowptr) (void)hipHostFree(rowptr);
        if (col) (void)hipHostFree(col);
    } else {
        free(rowptr);
        free(col);
    }
}

lt_status graph_alloc(Graph *g, uint32_t V, uint64_t E) {
    g->V = V;
    g->E = E;
    const size_t rb = sizeof(uint64_t) * ((size_t)V + 1);
    const size_t cb = sizeof(uint32_t) * (E ? E : 1);
    if (device_present()) {
        // Read-only after load: non-coherent lets the GPU cache zero-copy reads.
        const unsigned flags = hipHostMallocMapped | hipHostMallocPortable | hipHostMallocNonCoherent;
        if (hipHostMalloc((void **)&g->rowptr, rb, flags) != hipSuccess ||
            hipHostMalloc((void **)&g->col, cb, flags) != hipSuccess) {
            (void)hipGetLastError();
            set_error("pinned host allocation of %zu bytes failed", rb + cb);
            return LT_E_NOMEM;
        }
        g->pinned = true;
    } else {
        g->rowptr = (uint64_t *)malloc(rb);
        g->col = (uint32_t *)malloc(cb);
        if (!g->rowptr || !g->col) {
            set_error("host allocation of %zu bytes failed", rb + cb);
            return LT_E_NOMEM;
        }
    }
    return LT_OK;
}

// Binary search of PartitionStrategy::findPartition (PartitionStrategy.h:151-168).
uint32_t Partition::find(uint32_t v) const {
    uint32_t low = 0, high = P() - 1;
    while (low < high) {
        uint32_t mid = low + (high - low + 1) / 2;
        if (v < node_off[mid]) high = mid - 1;
        else low = mid;
    }
    return low;
}


// Partition table p describes graph g: it covers [0, V) and [0, E) with
// non-decreasing offsets, each edge offset is the rowptr of its first vertex,
// and every partition fits the partition size.  The engine and the cache
// sizing read slices of g through these offsets without further checks.
lt_status check_partition_of(const Graph &g, const Partition &p) {
    const size_t n = p.node_off.size();
    if (n < 2 || p.edge_off.size() != n) {
        set_error("partition table holds no partition");
        return LT_E_FORMAT;
    }
    const uint32_t P = (uint32_t)n - 1;
    if (p.node_off[0] != 0 || p.node_off[P] != g.V || p.edge_off[P] != g.E) {
        set_error("partition covers %u vertices / %llu edges, graph has %u / %llu", p.node_off[P],
                  (unsigned long long)p.edge_off[P], g.V, (unsigned long long)g.E);
        return LT_E_FORMAT;
    }
    for (uint32_t i = 0; i < P; i++)
        if (p.node_off[i + 1] < p.node_off[i]) {
            set_error("partition offsets decrease at partition %u", i);
            return LT_E_FORMAT;
        }
    for (uint32_t i = 0; i <= P; i++)
        if (p.edge_off[i] != g.rowptr[p.node_off[i]]) {
            set_error("partition %u edge offset does not match the graph", i);
            return LT_E_FORMAT;
        }
    for (uint32_t i = 0; i < P; i++) {
        const uint64_t need =
            ((uint64_t)p.node_off[i + 1] - p.node_off[i] + 1) * 8 + (p.edge_off[i + 1] - p.edge_off[i]) * 4;
        if (need > p.bytes) {
            set_error("partition %u needs %llu bytes > partition size %llu", i, (unsigned long long)need,
                      (unsigned long long)p.bytes);
            return LT_E_FORMAT;
        }
    }
    return LT_OK;
}

}  // namespace lt

using namespace lt;

extern "C" {

const char *lt_status_str(lt_status s) {
    switch (s) {
        case LT_OK: return "ok";
        case LT_E_IO: return "i/o error";
        case LT_E_FORMAT: return "bad file format";
        case LT_E_OVERSIZE_VERTEX: return "vertex larger than partition size";
        case LT_E_NOMEM: return "out of memory";
        case LT_E_HIP: return "HIP runtime error";
        case LT_E_ARG: return "invalid argument";
        case LT_E_NODEVICE: return "no HIP device";
        case LT_E_INTERNAL: return "internal error";
    }
    return "unknown status";
}

const char *lt_last_error(void) { return g_err; }

void lt_config_default(lt_config *c) {
    memset(c, 0, sizeof(*c));
    c->num_device_partitions = 0;
    c->num_device_walkers = 0;
    c->zerocopy_threshold = UINT64_MAX;
    c->seed = 0x4C54524Bull;  // "LTRK"
    c->policy = LT_SCHED_GREEDY;
    c->pipeline = 1;
    c->span = LT_SPAN_PARTITION;
    c->event_timers = 0;
    c->record_trace = 0;
    c->prefetch_depth = 2;
    c->prefetch_stick_pct = 50;
    c->evict_shortlist = 0;
}

// ---- bcsr (src/GraphLoader.h:30-48; tools/convert_from_txt.cc:211-218) ----
// A CSR the walk kernels may read without bounds checks: rowptr starts at 0,
// never decreases and ends at E (so every rowptr[v] <= E), and every column
// is a vertex id below V.
static lt_status check_csr(uint32_t V, uint64_t E, const uint64_t *rowptr, const uint32_t *col, const char *what) {
    if (rowptr[0] != 0 || rowptr[V] != E) {
        set_error("%s: rowptr[0]=%llu rowptr[V]=%llu E=%llu", what, (unsigned long long)rowptr[0],
                  (unsigned long long)rowptr[V], (unsigned long long)E);
        return LT_E_FORMAT;
    }
    for (uint32_t v = 0; v < V; v++)
        if (rowptr[v + 1] < rowptr[v]) {
            set_error("%s: rowptr decreases at %u", what, v);
            return LT_E_FORMAT;
        }
    for (uint64_t e = 0; e < E; e++)
        if (col[e] >= V) {
            set_error("%s: col[%llu]=%u >= V", what, (unsigned long long)e, col[e]);
            return LT_E_FORMAT;
        }
    return LT_OK;
}

// Layout: u32 num_node | u64 num_edge | u64 rowptr[V+1] | u32 col[E], packed.
lt_status lt_graph_open(const char *path, lt_graph **out) {
    if (!path || !out) return LT_E_ARG;
    FILE *f = fopen(path, "rb");
    if (!f) {
        set_error("cannot open %s", path);
        return LT_E_IO;
    }
    uint32_t V;
    uint64_t E;
    if (fread(&V, 4, 1, f) != 1 || fread(&E, 8, 1, f) != 1) {
        fclose(f);
        set_error("%s: short header", path);
        return LT_E_FORMAT;
    }
    fseek(f, 0, SEEK_END);
    const long long size = ftell(f);
    const long long want = 12 + 8LL * ((long long)V + 1) + 4LL * (long long)E;
    if (size != want) {
        fclose(f);
        set_error("%s: size %lld, header implies %lld", path, size, want);
        return LT_E_FORMAT;
    }
    fseek(f, 12, SEEK_SET);
    std::unique_ptr<lt_graph> g(new lt_graph);
    lt_status st = graph_alloc(g.get(), V, E);
    if (st != LT_OK) {
        fclose(f);
        return st;
    }
    if (fread(g->rowptr, 8, (size_t)V + 1, f) != (size_t)V + 1 || (E && fread(g->col, 4, E, f) != E)) {
        fclose(f);
        set_error("%s: short read", path);
        return LT_E_IO;
    }
    fclose(f);
    st = check_csr(V, E, g->rowptr, g->col, path);
    if (st != LT_OK) return st;
    *out = g.release();
    return LT_OK;
}

lt_status lt_graph_from_csr(uint32_t V, uint64_t E, const uint64_t *rowptr, const uint32_t *col, lt_graph **out) {
    if (!rowptr || (E && !col) || !out) return LT_E_ARG;
    lt_status cs = check_csr(V, E, rowptr, col, "csr");
    if (cs != LT_OK) return cs;
    std::unique_ptr<lt_graph> g(new lt_graph);
    lt_status st = graph_alloc(g.get(), V, E);
    if (st != LT_OK) return st;
    memcpy(g->rowptr, rowptr, 8 * ((size_t)V + 1));
    if (E) memcpy(g->col, col, 4 * E);
    *out = g.release();
    return LT_OK;
}

lt_status lt_graph_save(const lt_graph *g, const char *path) {
    if (!g || !path) return LT_E_ARG;
    FILE *f = fopen(path, "wb");
    if (!f) {
        set_error("cannot create %s", path);
        return LT_E_IO;
    }
    bool ok = fwrite(&g->V, 4, 1, f) == 1 && fwrite(&g->E, 8, 1, f) == 1 &&
              fwrite(g->rowptr, 8, (size_t)g->V + 1, f) == (size_t)g->V + 1 &&
              (g->E == 0 || fwrite(g->col, 4, g->E, f) == g->E);
    ok = (fclose(f) == 0) && ok;
    if (!ok) {
        set_error("write to %s failed", path);
        return LT_E_IO;
    }
    return LT_OK;
}

uint32_t lt_graph_num_node(const lt_graph *g) { return g ? g->V : 0; }
uint64_t lt_graph_num_edge(const lt_graph *g) { return g ? g->E : 0; }
const uint64_t *lt_graph_rowptr(const lt_graph *g) { return g ? g->rowptr : nullptr; }
const uint32_t *lt_graph_col(const lt_graph *g) { return g ? g->col : nullptr; }
void lt_graph_free(lt_graph *g) { delete g; }

// ---- partitions ---------------------------------------------------------
// Greedy split of EqualSizeStrategy::init (PartitionStrategy.h:189-235): grow
// the current range [n, i) while (i-n+1)*8 + (rowptr[i]-rowptr[n])*4 <= S;
// when vertex i-1 overflows it, the next partition starts at i-1.
lt_status lt_partition_build(const lt_graph *g, uint64_t max_bytes, lt_partition **out) {
    if (!g || !out || max_bytes == 0) return LT_E_ARG;
    std::unique_ptr<lt_partition> p(new lt_partition);
    p->bytes = max_bytes;
    p->node_off.push_back(0);
    p->edge_off.push_back(0);
    const uint64_t *rp = g->rowptr;
    uint64_t n = 0;
    for (uint64_t i = 1; i <= g->V;) {
        const uint64_t bytes = (i - n + 1) * 8u + (rp[i] - rp[n]) * 4u;
        if (bytes > max_bytes) {
            if (i == n + 1) {
                set_error("Node %llu has degree %llu. It cannot fit into partition size %llu",
                          (unsigned long long)n, (unsigned long long)(rp[i] - rp[n]),
                          (unsigned long long)max_bytes);
                return LT_E_OVERSIZE_VERTEX;
            }
            n = i - 1;
            p->node_off.push_back((uint32_t)n);
            p->edge_off.push_back(rp[n]);
            if (p->node_off.size() > 65536) {
                set_error("more than 65535 partitions (u16 partition ids)");
                return LT_E_ARG;
            }
        } else {
            i++;
        }
    }
    p->node_off.push_back(g->V);
    p->edge_off.push_back(rp[g->V]);
    *out = p.release();
    return LT_OK;
}

static std::string part_path(const char *graph_path) { return std::string(graph_path) + ".part"; }

// .part: u16 P | u64 partition_bytes | u32 node_off[P+1] | u64 edge_off[P+1]
// (PartitionStrategy.h:325-337 writes it, :76-97 reads it).
lt_status lt_partition_save(const lt_partition *p, const char *graph_path) {
    if (!p || !graph_path) return LT_E_ARG;
    if (p->P() > 65535) {  // the count field is a u16
        set_error("%u partitions do not fit a .part file (at most 65535); use a larger partition size", p->P());
        return LT_E_ARG;
    }
    const std::string path = part_path(graph_path);
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) {
        set_error("cannot create %s", path.c_str());
        return LT_E_IO;
    }
    const uint16_t P = (uint16_t)p->P();
    bool ok = fwrite(&P, 2, 1, f) == 1 && fwrite(&p->bytes, 8, 1, f) == 1 &&
              fwrite(p->node_off.data(), 4, p->node_off.size(), f) == p->node_off.size() &&
              fwrite(p->edge_off.data(), 8, p->edge_off.size(), f) == p->edge_off.size();
    ok = (fclose(f) == 0) && ok;
    if (!ok) {
        set_error("write to %s failed", path.c_str());
        return LT_E_IO;
    }
    return LT_OK;
}

lt_status lt_partition_load(const char *graph_path, lt_partition **out) {
    if (!graph_path || !out) return LT_E_ARG;
    const std::string path = part_path(graph_path);
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) {
        set_error("cannot open %s (materialise it with -m)", path.c_str());
        return LT_E_IO;
    }
    uint16_t P;
    uint64_t bytes;
    std::unique_ptr<lt_partition> p(new lt_partition);
    bool ok = fread(&P, 2, 1, f) == 1 && fread(&bytes, 8, 1, f) == 1;
    if (ok) {
        p->bytes = bytes;
        p->node_off.resize((size_t)P + 1);
        p->edge_off.resize((size_t)P + 1);
        ok = fread(p->node_off.data(), 4, (size_t)P + 1, f) == (size_t)P + 1 &&
             fread(p->edge_off.data(), 8, (size_t)P + 1, f) == (size_t)P + 1;
    }
    fclose(f);
    if (!ok || P == 0) {
        set_error("%s: short or empty partition file", path.c_str());
        return LT_E_FORMAT;
    }
    for (uint32_t i = 0; i < P; i++)
        if (p->node_off[i + 1] < p->node_off[i] || p->edge_off[i + 1] < p->edge_off[i]) {
            set_error("%s: offsets decrease at partition %u", path.c_str(), i);
            return LT_E_FORMAT;
        }
    *out = p.release();
    return LT_OK;
}

lt_status lt_partition_from_offsets(uint32_t P, uint64_t bytes, const uint32_t *node_off, const uint64_t *edge_off,
                                    lt_partition **out) {
    if (!node_off || !edge_off || !out || P == 0 || P > 65535) return LT_E_ARG;
    std::unique_ptr<lt_partition> p(new lt_partition);
    p->bytes = bytes;
    p->node_off.assign(node_off, node_off + P + 1);
    p->edge_off.assign(edge_off, edge_off + P + 1);
    for (uint32_t i = 0; i < P; i++)
        if (p->node_off[i + 1] < p->node_off[i] || p->edge_off[i + 1] < p->edge_off[i]) return LT_E_FORMAT;
    *out = p.release();
    return LT_OK;
}

uint32_t lt_partition_count(const lt_partition *p) { return p ? p->P() : 0; }
uint64_t lt_partition_bytes(const lt_partition *p) { return p ? p->bytes : 0; }
const uint32_t *lt_partition_node_offsets(const lt_partition *p) { return p ? p->node_off.data() : nullptr; }
const uint64_t *lt_partition_edge_offsets(const lt_partition *p) { return p ? p->edge_off.data() : nullptr; }
uint32_t lt_partition_find(const lt_partition *p, uint32_t v) { return p->find(v); }
void lt_partition_free(lt_partition *p) { delete p; }

// ---- pure scheduler rules -------------------------------------------------
// Scheduler::nextPartition, greedy (src/Scheduler.h:368-388): the partition
// with the most walkers, skipping the previously computed one unless the
// scheduler has just synchronised; strict '<' keeps the lowest id on ties.
int32_t lt_sched_greedy_next(uint32_t P, const uint64_t *counts, int32_t prev, int32_t sync) {
    uint64_t best = 0;
    int32_t choice = -1;
    for (uint32_t i = 0; i < P; i++) {
        if ((int32_t)i == prev && !sync) continue;
        if (best < counts[i]) {
            best = counts[i];
            choice = (int32_t)i;
        }
    }
    return choice;
}

// findPartitionwithleastWalker (src/Scheduler.h:149-171).  The reference walks
// an unordered_map; here ties resolve to the lowest partition id.
int32_t lt_sched_evict_min(uint32_t P, const uint64_t *counts, const uint8_t *resident, int32_t except) {
    int32_t victim = -1;
    uint64_t fewest = 0;
    for (uint32_t i = 0; i < P; i++) {
        if (!resident[i] || (int32_t)i == except) continue;
        if (victim < 0 || counts[i] < fewest) {
            fewest = counts[i];
            victim = (int32_t)i;
        }
    }
    return victim;
}

// getComputedPartition, greedy branch (src/Scheduler.h:84-111).
int32_t lt_sched_computed_partition(uint32_t P, const uint64_t *counts, const uint8_t *resident, int32_t except,
                                    uint64_t page) {
    int32_t full = -1, partial = -1;
    uint64_t full_min = 0, partial_max = 0;
    for (uint32_t i = 0; i < P; i++) {
        if ((int32_t)i == except || !resident[i]) continue;
        const uint64_t n = counts[i];
        if (n >= page) {
            if (full < 0 || n < full_min) {
                full = (int32_t)i;
                full_min = n;
            }
        } else if (full < 0 && n > 0 && (partial < 0 || n > partial_max)) {
            partial = (int32_t)i;
            partial_max = n;
        }
    }
    return full >= 0 ? full : partial;
}

// Belady's MIN (no counterpart in the reference): replay a demand sequence
// through a cache of `slots` slots, evicting on a miss the resident
// partition whose next demand is farthest ahead (never demanded again counts
// as infinitely far; ties go to the lowest id).  For unit-cost loads no
// eviction rule loads less on the same sequence, so the result bounds what
// any victim choice of the scheduler could save on a recorded trace.
lt_status lt_sched_belady(const int32_t *seq, uint64_t n, uint32_t P, uint32_t slots, const uint8_t *init_resident,
                          const uint64_t *part_bytes, uint64_t *loads_out, uint64_t *bytes_out, int32_t *victims_out) {
    if ((!seq && n) || P == 0 || slots == 0 || !loads_out) {
        set_error("lt_sched_belady: null argument, no partition or no slot");
        return LT_E_ARG;
    }
    for (uint64_t i = 0; i < n; i++)
        if (seq[i] < 0 || (uint32_t)seq[i] >= P) {
            set_error("lt_sched_belady: demand %llu names partition %d of %u", (unsigned long long)i, seq[i], P);
            return LT_E_ARG;
        }
    const uint64_t never = UINT64_MAX;
    std::vector<uint64_t> next_use(n), upcoming(P, never);  // next demand after position i; of each partition
    for (uint64_t i = n; i-- > 0;) {
        next_use[i] = upcoming[seq[i]];
        upcoming[seq[i]] = i;
    }
    // resident partitions ordered by their next demand, farthest first
    std::set<std::pair<uint64_t, int32_t>, std::greater<std::pair<uint64_t, int32_t>>> by_next;
    std::vector<uint64_t> key(P, 0);
    std::vector<uint8_t> in(P, 0);
    uint32_t used = 0;
    if (init_resident) {
        for (uint32_t p = 0; p < P && used < slots; p++)
            if (init_resident[p]) {
                in[p] = 1;
                key[p] = upcoming[p];
                by_next.insert({key[p], -(int32_t)p});  // negated id: the lowest id wins ties
                used++;
            }
    }
    uint64_t loads = 0, bytes = 0;
    for (uint64_t i = 0; i < n; i++) {
        const int32_t p = seq[i];
        int32_t victim = -1;
        if (in[p]) {
            by_next.erase({key[p], -p});
        } else {
            loads++;
            if (part_bytes) bytes += part_bytes[p];
            if (used == slots) {
                auto it = by_next.begin();
                victim = -it->second;
                by_next.erase(it);
                in[victim] = 0;
            } else {
                used++;
            }
            in[p] = 1;
        }
        key[p] = next_use[i];
        by_next.insert({key[p], -p});
        if (victims_out) victims_out[i] = victim;
    }
    *loads_out = loads;
    if (bytes_out) *bytes_out = bytes;
    return LT_OK;
}

lt_status lt_make_ppr_queries(const lt_graph *g, uint32_t nq, uint64_t lo, uint64_t hi, uint64_t seed,
                              uint32_t *sources_out, uint64_t *counts_out) {
    if (!g || !sources_out || !counts_out || g->E == 0 || lo == 0 || hi < lo) return LT_E_ARG;
    const double llo = log((double)lo), lhi = log((double)hi);
    for (uint32_t q = 0; q < nq; q++) {
        const lt::u32x4 x = lt::draw(seed, q, 0, lt::TAG_QUERY);
        const uint64_t e = lt::bounded(x.x, x.y, g->E);
        // last row r with rowptr[r] <= e
        const uint64_t *it = std::upper_bound(g->rowptr, g->rowptr + g->V + 1, e);
        sources_out[q] = (uint32_t)(it - g->rowptr - 1);
        const double u = (double)x.z / 4294967296.0;
        uint64_t c = (uint64_t)floor(exp(llo + u * (lhi - llo)));
        counts_out[q] = std::min<uint64_t>(std::max<uint64_t>(c, lo), hi);
    }
    return LT_OK;
}

lt_status lt_device_info(char *name, size_t cap, int *num_cu, uint64_t *hbm_bytes) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        (void)hipGetLastError();
        set_error("no HIP device visible");
        return LT_E_NODEVICE;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return LT_E_HIP;
    if (name && cap) {
        snprintf(name, cap, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (num_cu) *num_cu = prop.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = prop.totalGlobalMem;
    return LT_OK;
}

}  // extern "C"
