"""Packed partition copies on the GPU: unpack_kernel writes exactly the bytes
of a raw copy (the rowptr and col slices of GPUGraph::loadAsync,
src/GraphLoader.h:193-215) into an HBM slot, for sorted R-MAT graphs and for
rows stored absolute, and the engine's byte counters report what crossed
PCIe.  Walk parity with packing on is covered by the packed-* cases of
test_gpu_parity.py."""
import numpy as np
import pytest

import lighttraffic_amd as lt

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def slots_equal_raw(g, part, **kw):
    eng = lt.Engine(g, part, **kw)
    rowptr, col = g.rowptr(), g.col()
    no, eo = part.node_offsets(), part.edge_offsets()
    for p in range(part.count):
        rp, c = eng.fetch_partition(p, no, eo)
        assert np.array_equal(rp, rowptr[no[p]:no[p + 1] + 1]), p
        assert np.array_equal(c, col[eo[p]:eo[p + 1]]), p
    eng.close()


@pytest.mark.parametrize("scale,ef,pbytes", [(12, 8, 16 << 10), (16, 16, 1 << 20), (18, 16, 32 << 20)])
def test_packed_slots_equal_raw_rmat(scale, ef, pbytes):
    log(f"packed slots rmat s{scale} ef{ef} S={pbytes}")
    g = lt.Graph.rmat(scale, ef)
    part = lt.Partition.build(g, pbytes)
    slots_equal_raw(g, part, pack_partitions=1, num_device_partitions=2)
    slots_equal_raw(g, part, pack_partitions=1, num_device_partitions=2, rowptr_resident=1)
    slots_equal_raw(g, part, pack_partitions=0, num_device_partitions=2)


def test_packed_slots_unsorted_duplicates_empty_rows():
    log("packed slots: unsorted rows, duplicates, empty-row runs, degree-1 runs, hub row")
    rng = np.random.default_rng(11)
    V = 20000
    deg = rng.integers(0, 40, V)
    deg[rng.random(V) < 0.2] = 0
    deg[123] = 5000  # a row spanning many 64-edge chunks
    deg[1000:4000] = 0  # a long run of empty rows inside one chunk
    deg[5000:6000] = 1  # 64 rows ending inside one chunk
    rowptr = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint64)
    col = np.empty(int(rowptr[-1]), np.uint32)
    for r in range(V):
        a, b = int(rowptr[r]), int(rowptr[r + 1])
        x = np.sort(rng.integers(0, V, b - a))
        if r % 5 == 0:
            rng.shuffle(x)
        col[a:b] = x
    g = lt.Graph.from_csr(rowptr, col)
    part = lt.Partition.build(g, 64 << 10)
    assert part.count > 4
    slots_equal_raw(g, part, pack_partitions=1, num_device_partitions=3)


def test_packed_copies_with_edgeless_partitions():
    """Partitions holding only vertices without neighbours are copied raw (a
    rowptr slice, no columns) beside packed ones; slots and walks are exact."""
    from oracle import ffi as oracle
    log("packed copies with edgeless partitions")
    rng = np.random.default_rng(3)
    V = 9000
    deg = rng.integers(1, 20, V)
    deg[2000:7000] = 0
    rowptr = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint64)
    col = np.concatenate([np.sort(rng.integers(0, V, d)) for d in deg]).astype(np.uint32)
    g = lt.Graph.from_csr(rowptr, col)
    part = lt.Partition.build(g, 8192)
    eo = part.edge_offsets()
    assert sum(eo[p + 1] == eo[p] for p in range(part.count)) >= 2
    slots_equal_raw(g, part, pack_partitions=1, num_device_partitions=3)
    eng = lt.Engine(g, part, pack_partitions=1, num_device_partitions=3, zerocopy_threshold=0)
    w = 2 * V
    counts = eng.pagerank(w, 10)
    rc, _, _ = oracle.pagerank(rowptr, col, w, 10)
    assert np.array_equal(counts, rc)
    ends, moves = eng.genericwalk(w, 10, moves=True)
    re, rm, _ = oracle.genericwalk(rowptr, col, w, 10)
    assert np.array_equal(ends, re) and np.array_equal(moves, rm)


def test_drop_packed_images_keeps_running_engines():
    log("drop packed images")
    g = lt.Graph.rmat(12, 8)
    part = lt.Partition.build(g, 16 << 10)
    w = 2 * g.num_node
    a = lt.Engine(g, part, pack_partitions=1, num_device_partitions=3, zerocopy_threshold=0, pipeline=0)
    ref = a.pagerank(w, 10)
    first = a.stats()["packed_image_bytes"]  # a cold cache: every partition is copied at least once
    assert first > 0
    g.drop_packed()  # a keeps its image
    assert np.array_equal(a.pagerank(w, 10), ref) and a.stats()["packed_image_bytes"] > 0
    b = lt.Engine(g, part, pack_partitions=1, num_device_partitions=3, zerocopy_threshold=0,
                  pipeline=0)  # builds a new image
    assert np.array_equal(b.pagerank(w, 10), ref)
    assert b.stats()["packed_image_bytes"] == first  # same decisions, same images
    a.close()
    b.close()


def test_packed_byte_counters():
    log("packed byte counters")
    g = lt.Graph.rmat(14, 16)
    part = lt.Partition.build(g, 256 << 10)
    kw = dict(num_device_partitions=part.count // 3, pipeline=0, zerocopy_threshold=0)
    raw = lt.Engine(g, part, **kw)
    packed = lt.Engine(g, part, pack_partitions=1, **kw)
    w = 2 * g.num_node
    a = raw.pagerank(w, 10)
    b = packed.pagerank(w, 10)
    assert np.array_equal(a, b)
    sa, sb = raw.stats(), packed.stats()
    assert sa["graph_loading_calls"] == sb["graph_loading_calls"] > part.count
    assert sa["h2d_bytes_unpacked"] == sa["h2d_bytes_explicit"]
    assert sb["h2d_bytes_unpacked"] == sa["h2d_bytes_explicit"]
    assert sb["h2d_bytes_explicit"] < 0.75 * sa["h2d_bytes_explicit"]
    assert sa["packed_image_bytes"] == sa["packed_col_bytes"] == 0
    assert 0 < sb["packed_image_bytes"] < sb["packed_col_bytes"]
    assert sb["h2d_bytes_unpacked"] - sb["h2d_bytes_explicit"] == sb["packed_col_bytes"] - sb["packed_image_bytes"]
    log(f"explicit bytes raw {sa['h2d_bytes_explicit']} packed {sb['h2d_bytes_explicit']}")


def test_block_slots_hold_more_partitions():
    """pack_partitions=2 keeps 16-edge blocks packed in the slots: the same
    cache bytes hold more partitions, fewer partition loads are needed, and
    every walk equals the oracle's."""
    from oracle import ffi as oracle
    log("block slots")
    g = lt.Graph.rmat(14, 16)
    part = lt.Partition.build(g, 256 << 10)
    kw = dict(num_device_partitions=part.count // 3, pipeline=0, zerocopy_threshold=0)
    raw = lt.Engine(g, part, **kw)
    blk = lt.Engine(g, part, pack_partitions=2, **kw)
    w = 2 * g.num_node
    rc, re, _ = oracle.pagerank(g.rowptr(), g.col(), w, 10, endpoints=True)
    for eng in (raw, blk):
        c, e = eng.pagerank(w, 10, endpoints=True)
        assert np.array_equal(c, rc) and np.array_equal(e, re)
    sa, sb = raw.stats(), blk.stats()
    log(f"slots raw {sa['cache_slots']} blocks {sb['cache_slots']}; loads {sa['graph_loading_calls']} "
        f"{sb['graph_loading_calls']}; explicit bytes {sa['h2d_bytes_explicit']} {sb['h2d_bytes_explicit']}")
    assert sb["cache_slots"] > 1.3 * sa["cache_slots"]
    # the host-side layout query gives the engines' slot counts
    nd = kw["num_device_partitions"]
    assert part.slot_layout(g, 0, nd)[1] == sa["cache_slots"] == nd
    assert part.slot_layout(g, 2, nd)[1] == sb["cache_slots"]
    assert sb["graph_loading_calls"] < sa["graph_loading_calls"]
    assert 0 < sb["packed_image_bytes"] < 0.8 * sb["packed_col_bytes"]
    ends, moves = blk.genericwalk(w, 10, moves=True)
    re, rm, _ = oracle.genericwalk(g.rowptr(), g.col(), w, 10)
    assert np.array_equal(ends, re) and np.array_equal(moves, rm)
    counts, ends = blk.ppr(w, source=3, endpoints=True)
    rc, re, _ = oracle.ppr(g.rowptr(), g.col(), w, source=3, endpoints=True)
    assert np.array_equal(counts, rc) and np.array_equal(ends, re)
    with pytest.raises(lt.LTError):
        blk.fetch_partition(0, part.node_offsets(), part.edge_offsets())
    raw.close()
    blk.close()


def test_block_slots_edgeless_and_unsorted_partitions():
    from oracle import ffi as oracle
    log("block slots: edgeless partitions, unsorted rows, hub row")
    rng = np.random.default_rng(4)
    V = 20000
    deg = rng.integers(0, 40, V)
    deg[123] = 5000
    deg[1000:13000] = 0  # edgeless partitions (copied as a rowptr slice only)
    rowptr = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint64)
    col = np.empty(int(rowptr[-1]), np.uint32)
    for r in range(V):
        a, b = int(rowptr[r]), int(rowptr[r + 1])
        x = np.sort(rng.integers(0, V, b - a))
        if r % 5 == 0:
            rng.shuffle(x)
        col[a:b] = x
    g = lt.Graph.from_csr(rowptr, col)
    part = lt.Partition.build(g, 32 << 10)
    eo = part.edge_offsets()
    assert sum(eo[p + 1] == eo[p] for p in range(part.count)) >= 2
    eng = lt.Engine(g, part, pack_partitions=2, num_device_partitions=4, zerocopy_threshold=0)
    w = 2 * V
    counts, ends = eng.pagerank(w, 10, endpoints=True)
    rc, re, _ = oracle.pagerank(rowptr, col, w, 10, endpoints=True)
    assert np.array_equal(counts, rc) and np.array_equal(ends, re)
    ends, moves = eng.genericwalk(w, 10, moves=True)
    re, rm, _ = oracle.genericwalk(rowptr, col, w, 10)
    assert np.array_equal(ends, re) and np.array_equal(moves, rm)
    assert eng.stats()["packed_image_bytes"] > 0


CHILD = r"""
import sys
import numpy as np
import lighttraffic_amd as lt
from oracle import ffi as oracle
g = lt.Graph.rmat(13, 8)
part = lt.Partition.build(g, 16 << 10)
no, eo = part.node_offsets(), part.edge_offsets()
w = 2 * g.num_node
for kw in (dict(num_device_partitions=3, zerocopy_threshold=0),
           dict(num_device_partitions=4, span="resident", policy="learned", zerocopy_threshold=50)):
    eng = lt.Engine(g, part, pack_partitions=1, **kw)
    counts = eng.pagerank(w, 10)
    rc, _, _ = oracle.pagerank(g.rowptr(), g.col(), w, 10)
    assert np.array_equal(counts, rc), kw
    st = eng.stats()
    assert st["packed_image_bytes"] > 0 and st["graph_loading_calls"] > part.count, st
    ends, moves = eng.genericwalk(w, 10, moves=True)
    re, rm, _ = oracle.genericwalk(g.rowptr(), g.col(), w, 10)
    assert np.array_equal(ends, re) and np.array_equal(moves, rm), kw
    for p in range(part.count):
        rp, c = eng.fetch_partition(p, no, eo)
        assert np.array_equal(c, g.col()[eo[p]:eo[p + 1]]), p
    eng.close()
# block slots under the same queue setting (the bench's): no DMA stream, results exact
eng = lt.Engine(g, part, pack_partitions=2, num_device_partitions=4, span="resident", policy="learned",
                zerocopy_threshold=50)
rc, re, _ = oracle.pagerank(g.rowptr(), g.col(), w, 10, endpoints=True)
counts, ends = eng.pagerank(w, 10, endpoints=True)
assert np.array_equal(counts, rc) and np.array_equal(ends, re)
assert eng.stats()["packed_image_bytes"] > 0
eng.close()
print("child ok", part.count)
"""


def test_packed_copies_on_their_own_dma_stream():
    """With GPU_MAX_HW_QUEUES >= 6 (read when the HIP runtime starts, so in a
    child process) the packed copies' DMAs run on a stream of their own and
    two staging buffers alternate; results stay bit-exact."""
    import os
    import subprocess
    import sys
    log("packed copies with a DMA stream (child process, 8 hardware queues)")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, GPU_MAX_HW_QUEUES="8", PYTHONPATH=root)
    r = subprocess.run([sys.executable, "-c", CHILD], env=env, cwd=root, capture_output=True, text=True,
                       timeout=150)
    log(r.stdout.strip())
    assert r.returncode == 0, r.stderr[-2000:]
    assert "child ok" in r.stdout


def test_bandwidth_probes():
    """The two rates bench.py reports beside its measurements: pinned host ->
    HBM DMA (bounded by the PCIe link) and an HBM copy kernel."""
    log("bandwidth probes")
    h2d = lt.device_h2d_bandwidth(64 << 20, reps=2)
    hbm = lt.device_copy_bandwidth(256 << 20, reps=2)
    log(f"h2d {h2d:.1f} GB/s, hbm copy {hbm:.0f} GB/s")
    assert 5 < h2d < 130
    assert 500 < hbm < 16000
