// lt_internal.h — host-side objects behind the opaque handles of lighttraffic.h.
#pragma once
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/lighttraffic.h"

namespace lt {

void set_error(const char *fmt, ...);

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

}  // namespace lt

struct lt_graph : lt::Graph {};
struct lt_partition : lt::Partition {};
