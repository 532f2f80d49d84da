// lt_pack.cpp — packed partition images for the host->HBM partition copy.
//
// The reference copies a partition as its raw rowptr and col slices
// (GPUGraph::loadAsync, src/GraphLoader.h:193-215).  Here the col slice may
// instead travel gap-coded (format in lt_internal.h) and be unpacked in HBM by
// unpack_kernel into exactly the bytes a raw copy would have written, so the
// walk reads the same slot either way.  Sorted adjacency lists (what the bcsr
// converter and the R-MAT generator write) have small gaps between
// neighbours; a block holding a descending pair is stored absolute.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>

#include "lt_internal.h"

namespace lt {

static inline uint32_t bitlen(uint32_t x) { return x ? 32u - (uint32_t)__builtin_clz(x) : 0u; }

uint32_t pack_value_bits(uint32_t V) { return V <= 2 ? 1u : bitlen(V - 1); }

// Edge e (partition-local) of a row starting at row_begin is an anchor if it
// starts its row or its chunk, or its block is stored absolute.
static inline bool is_anchor(uint64_t e, uint64_t row_begin, uint8_t block_code) {
    return e == row_begin || e % kPackChunk == 0 || block_code == kPackRaw;
}

uint64_t pack_encode(const Graph &g, uint32_t nb, uint32_t ne, uint32_t vbits, uint8_t *out) {
    const uint64_t n = ne - nb;
    const uint64_t eb = g.rowptr[nb], m = g.rowptr[ne] - eb;
    const uint32_t *c = g.col + eb;
    const uint64_t nch = (m + kPackChunk - 1) / kPackChunk;
    const uint64_t nblk = nch * (kPackChunk / kPackBlock);
    // block widths: the widest gap between non-anchor neighbours of the block,
    // kPackRaw if one of those gaps is negative
    std::vector<uint8_t> code(nblk, 0);
    for (uint64_t r = 0; r < n; r++) {
        const uint64_t a = g.rowptr[nb + r] - eb, b = g.rowptr[nb + r + 1] - eb;
        for (uint64_t e = a + 1; e < b; e++) {
            if (e % kPackChunk == 0) continue;
            uint8_t &k = code[e / kPackBlock];
            if (k == kPackRaw) continue;
            if (c[e] < c[e - 1]) k = kPackRaw;
            else k = (uint8_t)std::max<uint32_t>(k, bitlen(c[e] - c[e - 1]));
        }
    }
    std::vector<uint64_t> chunk_bits(nch, 0);
    for (uint64_t r = 0; r < n; r++) {
        const uint64_t a = g.rowptr[nb + r] - eb, b = g.rowptr[nb + r + 1] - eb;
        for (uint64_t e = a; e < b; e++) {
            const uint8_t k = code[e / kPackBlock];
            chunk_bits[e / kPackChunk] += is_anchor(e, a, k) ? vbits : k;
        }
    }
    uint64_t words = 0;
    std::vector<uint64_t> chunk_word(nch);
    for (uint64_t k = 0; k < nch; k++) {
        chunk_word[k] = words;
        words += (chunk_bits[k] + 31) / 32;
    }
    if (words > UINT32_MAX) return UINT64_MAX;  // chunk_word would not fit 32 bits
    const uint64_t bytes = 8 * nch + 4 * (words + 2);
    if (!out) return bytes;
    uint32_t *meta = (uint32_t *)out;
    for (uint64_t k = 0; k < nch; k++) {
        meta[2 * k] = (uint32_t)chunk_word[k];
        uint32_t wc = 0;
        for (uint32_t j = 0; j < kPackChunk / kPackBlock; j++)
            wc |= (uint32_t)code[k * (kPackChunk / kPackBlock) + j] << (8 * j);
        meta[2 * k + 1] = wc;
    }
    uint32_t *wd = meta + 2 * nch;
    uint64_t pos = 0;  // bit position within the current chunk
    for (uint64_t r = 0; r < n; r++) {
        const uint64_t a = g.rowptr[nb + r] - eb, b = g.rowptr[nb + r + 1] - eb;
        for (uint64_t e = a; e < b; e++) {
            if (e % kPackChunk == 0) pos = 0;
            const uint8_t k = code[e / kPackBlock];
            const bool anchor = is_anchor(e, a, k);
            const uint32_t len = anchor ? vbits : k;
            const uint64_t v = anchor ? c[e] : (uint64_t)(c[e] - c[e - 1]);
            uint32_t *w = wd + chunk_word[e / kPackChunk] + pos / 32;
            const uint64_t sh = v << (pos % 32);
            w[0] |= (uint32_t)sh;
            if (pos % 32 + len > 32) w[1] |= (uint32_t)(sh >> 32);
            pos += len;
        }
    }
    return bytes;
}

uint64_t pack_blocks_encode(const Graph &g, uint32_t nb, uint32_t ne, uint8_t *out) {
    const uint64_t eb = g.rowptr[nb], m = g.rowptr[ne] - eb;
    const uint32_t *c = g.col + eb;
    const uint64_t nblk = (m + kBlockEdges - 1) / kBlockEdges;
    uint64_t bits = 0;
    std::vector<uint32_t> base(nblk), width(nblk);
    for (uint64_t k = 0; k < nblk; k++) {
        const uint64_t a = k * kBlockEdges, b = std::min<uint64_t>(a + kBlockEdges, m);
        uint32_t lo = c[a], hi = c[a];
        for (uint64_t e = a + 1; e < b; e++) {
            lo = std::min(lo, c[e]);
            hi = std::max(hi, c[e]);
        }
        base[k] = lo;
        width[k] = bitlen(hi - lo);
        bits += (uint64_t)width[k] * kBlockEdges;
    }
    if (bits > UINT32_MAX) return UINT64_MAX;
    const uint64_t words = (bits + 31) / 32;
    const uint64_t bytes = 8 * (nblk + 1) + 4 * (words + 2);
    if (!out) return bytes;
    uint32_t *hdr = (uint32_t *)out;
    uint32_t *wd = hdr + 2 * (nblk + 1);
    uint64_t pos = 0;
    for (uint64_t k = 0; k < nblk; k++) {
        hdr[2 * k] = base[k];
        hdr[2 * k + 1] = (uint32_t)pos;
        const uint64_t a = k * kBlockEdges;
        for (uint64_t j = 0; j < kBlockEdges; j++) {
            const uint64_t e = a + j;
            const uint64_t v = e < m ? (uint64_t)(c[e] - base[k]) : 0;  // padding: the base itself
            const uint64_t sh = v << (pos % 32);
            uint32_t *w = wd + pos / 32;
            w[0] |= (uint32_t)sh;
            if (pos % 32 + width[k] > 32) w[1] |= (uint32_t)(sh >> 32);
            pos += width[k];
        }
    }
    hdr[2 * nblk] = 0;
    hdr[2 * nblk + 1] = (uint32_t)pos;
    return bytes;
}

void slot_layout(const Partition &p, const uint64_t *sizes, uint32_t raw_slots, uint64_t *slot_bytes,
                 uint32_t *slots) {
    const uint32_t P = p.P();
    const uint64_t raw = (p.bytes + 255) & ~255ull;
    uint32_t n = raw_slots == 0 ? P : std::min(raw_slots, P);
    uint64_t sb = raw;
    if (sizes) {
        uint64_t need = 0;
        for (uint32_t i = 0; i < P; i++) {
            const uint64_t rb = ((uint64_t)p.node_off[i + 1] - p.node_off[i] + 1) * 8;
            need = std::max(need, rb + (sizes[i] ? sizes[i] : 4 * (p.edge_off[i + 1] - p.edge_off[i])));
        }
        sb = std::max<uint64_t>(256, (need + 255) & ~255ull);
        n = (uint32_t)std::min<uint64_t>(P, (uint64_t)n * raw / sb);
    }
    *slot_bytes = sb;
    *slots = n;
}

PackedImage::~PackedImage() {
    if (!host) return;
    if (pinned) (void)hipHostFree(host);
    else free(host);
}

// Runs f(i) for i in [0, n) on up to 16 threads.
template <class F>
static void parallel_for(uint32_t n, F f) {
    const uint32_t nt = std::max(1u, std::min({n, 16u, std::thread::hardware_concurrency()}));
    std::atomic<uint32_t> next{0};
    std::vector<std::thread> th;
    for (uint32_t t = 0; t < nt; t++)
        th.emplace_back([&] {
            for (uint32_t i; (i = next.fetch_add(1)) < n;) f(i);
        });
    for (auto &t : th) t.join();
}

lt_status packed_image(const Graph &g, const Partition &p, int format, std::shared_ptr<const PackedImage> *out) {
    for (auto &pk : g.packs)
        if (pk->format == format && pk->bytes == p.bytes && pk->node_off == p.node_off && pk->edge_off == p.edge_off) {
            *out = pk;
            return LT_OK;
        }
    auto pk = std::make_shared<PackedImage>();
    const uint32_t P = p.P();
    pk->format = format;
    pk->bytes = p.bytes;
    pk->node_off = p.node_off;
    pk->edge_off = p.edge_off;
    pk->vbits = pack_value_bits(g.V);
    pk->off.assign(P, 0);
    pk->size.assign(P, 0);
    auto encode = [&](uint32_t i, uint8_t *dst) {
        return format == 2 ? pack_blocks_encode(g, p.node_off[i], p.node_off[i + 1], dst)
                           : pack_encode(g, p.node_off[i], p.node_off[i + 1], pk->vbits, dst);
    };
    parallel_for(P, [&](uint32_t i) { pk->size[i] = encode(i, nullptr); });
    // a partition is copied packed only when that copies fewer bytes
    for (uint32_t i = 0; i < P; i++) {
        const uint64_t raw = 4 * (p.edge_off[i + 1] - p.edge_off[i]);
        if (pk->size[i] >= raw) {
            pk->size[i] = 0;
            continue;
        }
        pk->off[i] = pk->total;
        pk->total += (pk->size[i] + 255) & ~255ull;
        pk->max_size = std::max(pk->max_size, pk->size[i]);
    }
    if (pk->total) {
        if (g.pinned) {
            // mapped: the device-scheduled loop's copy kernel reads block images in place
            if (hipHostMalloc((void **)&pk->host, pk->total, hipHostMallocMapped | hipHostMallocPortable) !=
                hipSuccess) {
                (void)hipGetLastError();
                pk->host = nullptr;
                set_error("pinned host allocation of %llu bytes for packed partitions failed",
                          (unsigned long long)pk->total);
                return LT_E_NOMEM;
            }
            pk->pinned = true;
        } else if (!(pk->host = (uint8_t *)malloc(pk->total))) {
            set_error("host allocation of %llu bytes for packed partitions failed", (unsigned long long)pk->total);
            return LT_E_NOMEM;
        }
        parallel_for(P, [&](uint32_t i) {
            if (!pk->size[i]) return;
            uint8_t *dst = pk->host + pk->off[i];
            memset(dst, 0, pk->size[i]);
            encode(i, dst);
        });
    }
    *out = pk;
    g.packs.push_back(std::move(pk));
    return LT_OK;
}

}  // namespace lt

extern "C" {

void lt_graph_drop_packed(lt_graph *g) {
    if (g) g->packs.clear();
}

static lt_status pack_one(const char *fn, int format, const lt_graph *g, const lt_partition *p, uint32_t part,
                          void *out, uint64_t cap, uint64_t *bytes_out) {
    if (!g || !p || !bytes_out || part >= p->P() || p->node_off[part + 1] > g->V) {
        lt::set_error("%s: null handle, or partition %u not in the graph", fn, part);
        return LT_E_ARG;
    }
    const uint32_t nb = p->node_off[part], ne = p->node_off[part + 1];
    const uint32_t vbits = lt::pack_value_bits(g->V);
    auto encode = [&](uint8_t *dst) {
        return format == 2 ? lt::pack_blocks_encode(*g, nb, ne, dst) : lt::pack_encode(*g, nb, ne, vbits, dst);
    };
    const uint64_t bytes = encode(nullptr);
    if (bytes == UINT64_MAX) {
        lt::set_error("%s: partition %u too large to pack", fn, part);
        return LT_E_ARG;
    }
    *bytes_out = bytes;
    if (!out) return LT_OK;
    if (cap < bytes) {
        lt::set_error("%s: %llu bytes needed, %llu given", fn, (unsigned long long)bytes, (unsigned long long)cap);
        return LT_E_ARG;
    }
    memset(out, 0, bytes);
    encode((uint8_t *)out);
    return LT_OK;
}

lt_status lt_partition_pack(const lt_graph *g, const lt_partition *p, uint32_t part, void *out, uint64_t cap,
                            uint64_t *bytes_out) {
    return pack_one("lt_partition_pack", 1, g, p, part, out, cap, bytes_out);
}

lt_status lt_partition_slot_layout(const lt_graph *g, const lt_partition *p, int32_t pack_partitions,
                                   uint32_t num_device_partitions, uint64_t *slot_bytes, uint32_t *slots) {
    if (!g || !p || !slot_bytes || !slots || pack_partitions < 0 || pack_partitions > 2) {
        lt::set_error("lt_partition_slot_layout: null argument or pack_partitions not 0-2");
        return LT_E_ARG;
    }
    if (lt::check_partition_of(*g, *p) != LT_OK) return LT_E_ARG;  // the message says what does not match
    std::vector<uint64_t> sizes;
    bool any = false;
    if (pack_partitions == 2) {  // the sizes an engine's block image gets (lt_pack.cpp packed_image)
        const uint32_t P = p->P();
        sizes.assign(P, 0);
        lt::parallel_for(P, [&](uint32_t i) {
            const uint64_t s = lt::pack_blocks_encode(*g, p->node_off[i], p->node_off[i + 1], nullptr);
            sizes[i] = s < 4 * (p->edge_off[i + 1] - p->edge_off[i]) ? s : 0;
        });
        for (uint64_t s : sizes) any = any || s;
    }
    lt::slot_layout(*p, any ? sizes.data() : nullptr, num_device_partitions, slot_bytes, slots);
    return LT_OK;
}

lt_status lt_partition_pack_blocks(const lt_graph *g, const lt_partition *p, uint32_t part, void *out,
                                   uint64_t cap, uint64_t *bytes_out) {
    return pack_one("lt_partition_pack_blocks", 2, g, p, part, out, cap, bytes_out);
}

namespace {
// Bytes a block slot needs for a vertex range while its vertices are added in
// order: the rowptr slice plus the block image pack_blocks_encode writes for
// the range's columns, or the raw columns when that image is not smaller
// (the engine then copies them as they are; lt::slot_layout sizes slots the
// same way).  Blocks are 16 consecutive edges counted from the range's first.
struct BlockSizer {
    uint64_t nv = 0, ne = 0;    // vertices and edges added
    uint64_t nblk = 0, bits = 0;  // complete blocks and their stream bits
    uint32_t open = 0, lo = 0, hi = 0;  // edges of the open block and their range

    void add(const uint32_t *c, uint64_t d) {
        nv++;
        ne += d;
        for (uint64_t k = 0; k < d; k++) {
            const uint32_t x = c[k];
            if (open == 0) {
                lo = hi = x;
            } else {
                lo = std::min(lo, x);
                hi = std::max(hi, x);
            }
            if (++open == lt::kBlockEdges) {
                bits += (uint64_t)lt::kBlockEdges * lt::bitlen(hi - lo);
                nblk++;
                open = 0;
            }
        }
    }
    uint64_t slot_bytes() const {
        const uint64_t tb = bits + (open ? (uint64_t)lt::kBlockEdges * lt::bitlen(hi - lo) : 0);
        const uint64_t nb = nblk + (open ? 1 : 0);
        const uint64_t img = tb > UINT32_MAX ? UINT64_MAX : 8 * (nb + 1) + 4 * ((tb + 31) / 32 + 2);
        return (nv + 1) * 8 + std::min(img, 4 * ne);
    }
};
}  // namespace

lt_status lt_partition_build_blocks(const lt_graph *g, uint64_t slot_bytes, lt_partition **out) {
    if (!g || !out || slot_bytes == 0) {
        lt::set_error("lt_partition_build_blocks: null argument or zero slot size");
        return LT_E_ARG;
    }
    std::unique_ptr<lt_partition> p(new lt_partition);
    p->node_off.push_back(0);
    p->edge_off.push_back(0);
    const uint64_t *rp = g->rowptr;
    uint64_t raw_max = 0;
    uint32_t n = 0;  // first vertex of the open partition
    auto close = [&](uint32_t end) {
        raw_max = std::max<uint64_t>(raw_max, ((uint64_t)end - n + 1) * 8 + (rp[end] - rp[n]) * 4);
        p->node_off.push_back(end);
        p->edge_off.push_back(rp[end]);
    };
    BlockSizer cur;
    for (uint32_t v = 0; v < g->V; v++) {
        BlockSizer next = cur;
        next.add(g->col + rp[v], rp[v + 1] - rp[v]);
        if (next.slot_bytes() <= slot_bytes) {
            cur = next;
            continue;
        }
        // v does not fit the open partition: it starts the next one, where
        // it must fit alone (when v == n it was alone already)
        if (v > n) {
            close(v);
            n = v;
            cur = BlockSizer();
            cur.add(g->col + rp[v], rp[v + 1] - rp[v]);
        }
        if (v == n && (cur.nv == 0 || cur.slot_bytes() > slot_bytes)) {
            lt::set_error("Node %u has degree %llu. It cannot fit into a block slot of %llu bytes", v,
                          (unsigned long long)(rp[v + 1] - rp[v]), (unsigned long long)slot_bytes);
            return LT_E_OVERSIZE_VERTEX;
        }
        if (p->node_off.size() > 65535) {
            lt::set_error("more than 65535 partitions (u16 partition ids)");
            return LT_E_ARG;
        }
    }
    close(g->V);
    p->bytes = raw_max;
    *out = p.release();
    return LT_OK;
}

}  // extern "C"
