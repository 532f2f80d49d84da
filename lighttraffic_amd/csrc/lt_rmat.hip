// lt_rmat.hip — seeded R-MAT graph generator on the GPU.
//
// Produces the same bcsr as writing the sampled edges to a text edge list and
// running the reference converter tools/convert_from_txt.cc:
//   read_txt (:66-103)              drop self loops
//   remove_isolated_vertices (:105-138)  relabel ids densely by rank
//   to_undirected (:140-162)        add reverse edges, sort by (src,dst), dedupe
//   main (:188-218)                 V = last src + 1, rowptr by counting, col = dst
// Edge e, level t takes word t%4 of Philox(seed; e, t/4, TAG_RMAT) and picks a
// quadrant by the cumulative thresholds of (a, a+b, a+b+c); level t sets bit
// scale-1-t of (src, dst).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <stdlib.h>

#include <memory>

#include "lt_internal.h"
#include "lt_rng.h"

#define RM_HIP(call)                                                                                   \
    do {                                                                                               \
        hipError_t _e = (call);                                                                        \
        if (_e != hipSuccess) {                                                                        \
            lt::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(_e));            \
            return LT_E_HIP;                                                                           \
        }                                                                                              \
    } while (0)

namespace lt {

__global__ void rmat_sample_kernel(uint64_t n, uint32_t scale, uint32_t ta, uint32_t tab, uint32_t tabc, uint64_t seed,
                                   uint32_t *src, uint32_t *dst, uint32_t *present) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride) {
        uint32_t s = 0, d = 0;
        u32x4 r{0, 0, 0, 0};
        for (uint32_t t = 0; t < scale; t++) {
            if ((t & 3u) == 0) r = draw(seed, e, t >> 2, TAG_RMAT);
            const uint32_t u = (t & 3u) == 0 ? r.x : (t & 3u) == 1 ? r.y : (t & 3u) == 2 ? r.z : r.w;
            const uint32_t sb = u >= tab ? 1u : 0u;
            const uint32_t db = (u >= ta && u < tab) || u >= tabc ? 1u : 0u;
            s |= sb << (scale - 1 - t);
            d |= db << (scale - 1 - t);
        }
        src[e] = s;
        dst[e] = d;
        if (s != d) {
            present[s] = 1u;
            present[d] = 1u;
        }
    }
}

// Both directions of every non-loop edge as (newid(s) << scale | newid(d));
// self loops become the sentinel, which sorts last.
__global__ void rmat_keys_kernel(uint64_t n, uint32_t scale, const uint32_t *src, const uint32_t *dst,
                                 const uint32_t *newid, uint64_t *keys) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t sentinel = (scale == 32) ? ~0ull : ((1ull << (2 * scale)) - 1);
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride) {
        const uint32_t s = src[e], d = dst[e];
        if (s == d) {
            keys[2 * e] = sentinel;
            keys[2 * e + 1] = sentinel;
        } else {
            const uint64_t a = newid[s], b = newid[d];
            keys[2 * e] = (a << scale) | b;
            keys[2 * e + 1] = (b << scale) | a;
        }
    }
}

// Sorted unique keys -> rowptr (first edge of each row) and col.
__global__ void rmat_csr_kernel(uint64_t E, uint32_t scale, const uint64_t *keys, uint64_t *rowptr, uint32_t *col) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t mask = (scale == 32) ? 0xFFFFFFFFull : ((1ull << scale) - 1);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < E; i += stride) {
        const uint64_t k = keys[i];
        const uint64_t r = k >> scale;
        col[i] = (uint32_t)(k & mask);
        const uint64_t prev = i == 0 ? 0 : (keys[i - 1] >> scale) + 1;
        for (uint64_t x = prev; x <= r; x++) rowptr[x] = i;
    }
}

template <typename T>
struct DevBuf {
    T *p = nullptr;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
    }
    hipError_t alloc(size_t n) {
        release();
        return hipMalloc((void **)&p, sizeof(T) * (n ? n : 1));
    }
};

static lt_status generate(uint32_t scale, uint32_t ef, double a, double b, double c, uint64_t seed, lt_graph **out) {
    if (scale == 0 || scale > 31 || ef == 0 || !(a >= 0) || !(b >= 0) || !(c >= 0) || a + b + c > 1.0) {
        set_error("bad R-MAT parameters");
        return LT_E_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        set_error("the R-MAT generator needs a HIP device");
        return LT_E_NODEVICE;
    }
    const uint64_t n = (uint64_t)ef << scale;
    const uint64_t nv = 1ull << scale;
    hipStream_t s;
    RM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    std::unique_ptr<hipStream_t, void (*)(hipStream_t *)> sg(&s, [](hipStream_t *x) { (void)hipStreamDestroy(*x); });
    const int grid = 4096, block = 256;

    DevBuf<uint32_t> src, dst, present, newid;
    RM_HIP(src.alloc(n));
    RM_HIP(dst.alloc(n));
    RM_HIP(present.alloc(nv));
    RM_HIP(newid.alloc(nv + 1));
    RM_HIP(hipMemsetAsync(present.p, 0, 4 * nv, s));
    rmat_sample_kernel<<<grid, block, 0, s>>>(n, scale, prob_threshold(a), prob_threshold(a + b),
                                              prob_threshold(a + b + c), seed, src.p, dst.p, present.p);
    RM_HIP(hipGetLastError());

    // dense relabelling: newid = exclusive prefix count of present vertices
    size_t tmp_bytes = 0;
    RM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, present.p, newid.p, nv + 0, s));
    DevBuf<uint8_t> tmp;
    RM_HIP(tmp.alloc(tmp_bytes));
    RM_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tmp_bytes, present.p, newid.p, nv, s));
    present.release();

    const uint64_t nk = 2 * n;
    DevBuf<uint64_t> keys, alt;
    RM_HIP(keys.alloc(nk));
    rmat_keys_kernel<<<grid, block, 0, s>>>(n, scale, src.p, dst.p, newid.p, keys.p);
    RM_HIP(hipGetLastError());
    RM_HIP(hipStreamSynchronize(s));
    src.release();
    dst.release();
    newid.release();
    RM_HIP(alt.alloc(nk));

    hipcub::DoubleBuffer<uint64_t> db(keys.p, alt.p);
    const int end_bit = (int)(2 * scale);
    tmp_bytes = 0;
    RM_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, tmp_bytes, db, nk, 0, end_bit, s));
    RM_HIP(tmp.alloc(tmp_bytes));
    RM_HIP(hipcub::DeviceRadixSort::SortKeys(tmp.p, tmp_bytes, db, nk, 0, end_bit, s));
    uint64_t *sorted = db.Current();
    uint64_t *uniq = (sorted == keys.p) ? alt.p : keys.p;

    // Deduplicate in pieces of at most 2^30 keys.  Piece j > 0 starts one key
    // early: that key equals the last one already kept, so its copy lands on
    // the same slot and the piece adds nsel-1 new keys.
    DevBuf<uint64_t> d_nsel;
    RM_HIP(d_nsel.alloc(1));
    uint64_t piece = 1ull << 30;
    if (const char *env = getenv("LT_RMAT_DEDUPE_PIECE")) {  // smaller pieces exercise the seams in tests
        const uint64_t v = strtoull(env, nullptr, 10);
        if (v >= 2) piece = v;
    }
    uint64_t nsel = 0, last = 0;
    for (uint64_t lo = 0; lo < nk; lo += piece) {
        const uint64_t hi = nk - lo < piece ? nk : lo + piece;
        const uint64_t first = lo == 0 ? 0 : lo - 1;
        uint64_t *dst_ptr = uniq + (lo == 0 ? 0 : nsel - 1);
        tmp_bytes = 0;
        RM_HIP(hipcub::DeviceSelect::Unique(nullptr, tmp_bytes, sorted + first, dst_ptr, d_nsel.p, hi - first, s));
        RM_HIP(tmp.alloc(tmp_bytes));
        RM_HIP(hipcub::DeviceSelect::Unique(tmp.p, tmp_bytes, sorted + first, dst_ptr, d_nsel.p, hi - first, s));
        uint64_t got = 0;
        RM_HIP(hipMemcpyAsync(&got, d_nsel.p, 8, hipMemcpyDeviceToHost, s));
        RM_HIP(hipStreamSynchronize(s));
        nsel += lo == 0 ? got : got - 1;
    }
    const uint64_t sentinel = (1ull << (2 * scale)) - 1;
    uint64_t E = nsel;
    if (E) {
        RM_HIP(hipMemcpy(&last, uniq + E - 1, 8, hipMemcpyDeviceToHost));
        if (last == sentinel) {
            E--;
            if (E) RM_HIP(hipMemcpy(&last, uniq + E - 1, 8, hipMemcpyDeviceToHost));
        }
    }
    const uint32_t V = E ? (uint32_t)(last >> scale) + 1 : 0;
    if (V == 0) {
        set_error("R-MAT sample has no edges after removing self loops");
        return LT_E_ARG;
    }
    std::unique_ptr<lt_graph> g(new lt_graph);
    lt_status st = graph_alloc(g.get(), V, E);
    if (st != LT_OK) return st;
    DevBuf<uint64_t> rowptr;
    DevBuf<uint32_t> col;
    RM_HIP(rowptr.alloc((size_t)V + 1));
    RM_HIP(col.alloc(E));
    rmat_csr_kernel<<<grid, block, 0, s>>>(E, scale, uniq, rowptr.p, col.p);
    RM_HIP(hipGetLastError());
    RM_HIP(hipMemcpyAsync(rowptr.p + V, &E, 8, hipMemcpyHostToDevice, s));
    RM_HIP(hipMemcpyAsync(g->rowptr, rowptr.p, 8 * ((size_t)V + 1), hipMemcpyDeviceToHost, s));
    RM_HIP(hipMemcpyAsync(g->col, col.p, 4 * E, hipMemcpyDeviceToHost, s));
    RM_HIP(hipStreamSynchronize(s));
    *out = g.release();
    return LT_OK;
}

}  // namespace lt

extern "C" lt_status lt_graph_generate_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                                            uint64_t seed, lt_graph **out) {
    if (!out) return LT_E_ARG;
    return lt::generate(scale, edge_factor, a, b, c, seed, out);
}
