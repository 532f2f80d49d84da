// lt_internal.h — host-side objects behind the opaque handles of lighttraffic.h.
#pragma once
#include <stdint.h>
#include <memory>
#include <string>
#include <vector>

#include "../../include/lighttraffic.h"

namespace lt {

void set_error(const char *fmt, ...);

struct PackedImage;

// bcsr graph (src/GraphLoader.h:30-48).  rowptr/col live in pinned,
// device-mapped host memory when a HIP device is present, so that the engine
// can copy partitions with DMA or read them in place over PCIe; without a
// device they are ordinary heap memory (file tools and host tests).
struct Graph {
    uint32_t V = 0;
    uint64_t E = 0;
    uint64_t *rowptr = nullptr;
    uint32_t *col = nullptr;
    bool pinned = false;
    // Packed partition images built for this graph, one per partition table
    // (built on first use by an engine that copies packed partitions; an
    // engine shares ownership of the image it uses).
    mutable std::vector<std::shared_ptr<const PackedImage>> packs;
    ~Graph();
};

// Alloc/free of graph arrays: pinned + mapped when a device is available.
lt_status graph_alloc(Graph *g, uint32_t V, uint64_t E);

// Contiguous vertex-range partitions (src/partition/PartitionStrategy.h).
struct Partition {
    uint64_t bytes = 0;                // partition byte budget S
    std::vector<uint32_t> node_off;    // P+1
    std::vector<uint64_t> edge_off;    // P+1
    uint32_t P() const { return (uint32_t)node_off.size() - 1; }
    uint32_t find(uint32_t v) const;
};

// Partition table p describes graph g (coverage, offsets, sizes); LT_E_FORMAT otherwise.
lt_status check_partition_of(const Graph &g, const Partition &p);

// Packed partition image: the column slice of a partition, gap-coded for the
// host->HBM copy and unpacked in HBM by unpack_kernel.  Layout of one image
// for a partition of m edges (partition-local edge index e), nch = ceil(m/64)
// chunks of 64 edges, each of 4 blocks of 16 edges:
//   u32 meta[2*nch]: meta[2c] = first 32-bit word of chunk c in words[];
//       meta[2c+1] byte j = code of block j of chunk c: 0..32 = bit width of
//       its gaps, kPackRaw = block stored absolute;
//   u32 words[]: the chunks' bit streams, each starting on a word, LSB first,
//       followed by two zero words.
// Edge e is an anchor if it is the first edge of its row or of its chunk, or
// its block is kPackRaw; an anchor holds col[e] in vbits bits, any other edge
// holds col[e] - col[e-1] in its block's width.  So every chunk decodes
// without the others (the row starts come from the slot's rowptr), and a
// block is gap-coded only when its non-anchor edges are in non-decreasing
// order within their rows.
constexpr uint32_t kPackChunk = 64;
constexpr uint32_t kPackBlock = 16;
constexpr uint8_t kPackRaw = 0xFF;

// Block image (pack_partitions = 2): the column slice in a form the walk
// kernel reads in place, kept packed in the HBM slot.  For m edges in
// nblk = ceil(m/16) blocks of 16 (the last padded with its base):
//   u32x2 hdr[nblk + 1]: {base, bit offset} of block k; hdr[nblk] = {0, total bits};
//   u32 words[]: the blocks' bit streams back to back (LSB first), then two
//       zero words.
// Block k holds col[e] - base for its 16 edges in w = (off[k+1] - off[k]) / 16
// bits each (base = the block's smallest column), so column e is
// hdr[e/16].base + bits(off[e/16] + (e % 16) * w, w).
constexpr uint32_t kBlockEdges = 16;

struct PackedImage {
    int format = 1;                      // 1 gap-coded chunks, 2 blocks read in place
    uint64_t bytes = 0;                  // partition budget S the image is for
    std::vector<uint32_t> node_off;      // partition table it was built for
    std::vector<uint64_t> edge_off;
    std::vector<uint64_t> off, size;     // per partition: offset, bytes (0 = copy unpacked)
    uint8_t *host = nullptr;
    bool pinned = false;
    uint64_t total = 0, max_size = 0;
    uint32_t vbits = 1;
    ~PackedImage();
};

// Bits of an anchor: enough for V-1, at least 1.
uint32_t pack_value_bits(uint32_t V);
// Bytes of the image of rows [nb, ne), UINT64_MAX if its bit streams would
// exceed 2^32 words; writes it when out is not null (out must hold that many
// zeroed bytes).
uint64_t pack_encode(const Graph &g, uint32_t nb, uint32_t ne, uint32_t vbits, uint8_t *out);
// Bytes of the block image of rows [nb, ne), UINT64_MAX if its bit offsets
// would exceed 32 bits; writes it when out is not null (zeroed, that size).
uint64_t pack_blocks_encode(const Graph &g, uint32_t nb, uint32_t ne, uint8_t *out);
// Slot bytes (256-aligned) and number of slots of an HBM partition cache of
// `raw_slots` raw slots (0: one per partition) for partition table p.  sizes:
// per partition, the bytes of the packed column image its slot keeps (0: raw
// columns), or nullptr for raw slots; the cache keeps raw_slots raw slots'
// bytes, and holds at most one slot per partition.
void slot_layout(const Partition &p, const uint64_t *sizes, uint32_t raw_slots, uint64_t *slot_bytes,
                 uint32_t *slots);
// The graph's image for partition table p in `format` (1 or 2), built on
// first call.
lt_status packed_image(const Graph &g, const Partition &p, int format, std::shared_ptr<const PackedImage> *out);

}  // namespace lt

struct lt_graph : lt::Graph {};
struct lt_partition : lt::Partition {};
