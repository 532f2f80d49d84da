"""GPU parity: the HIP engine against the CPU oracle, bit for bit.

Every case runs the engine through the C ABI (liblighttraffic.so) and the
oracle on the same graph, seed and walker ids, and compares per-walker
endpoints / move counts and the u32 visit counters exactly.  Cases cover the
reference's three programs (app/pagerank.cu, app/ppr.cu, app/genericwalk.cu),
one or many partitions, a capped HBM cache with loads and evictions,
zero-copy reads, the FIFO variant, pipelining on and off, walkers processed in
several waves (-b), and vertices without neighbours.
"""
import os

import numpy as np
import pytest

import lighttraffic_amd as lt
from oracle import ffi as oracle

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SEED = 0x4C54524B


def log(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def g12():
    return lt.Graph.rmat(12, 8)


@pytest.mark.parametrize("name,scale,ef,seed", [("rmat_s8_ef4", 8, 4, 0x5EED0008), ("rmat_s10_ef8", 10, 8, 0x5EED000A)])
def test_generator_matches_reference_converter(name, scale, ef, seed, tmp_path):
    """GPU R-MAT pipeline == oracle edges run through tools/convert_from_txt.cc."""
    log(f"generator {name}")
    g = lt.Graph.rmat(scale, ef, seed=seed)
    out = tmp_path / "g.bcsr"
    g.save(str(out))
    with open(os.path.join(GOLDEN, name + ".bcsr"), "rb") as f:
        assert out.read_bytes() == f.read()


def test_generator_matches_oracle_s16():
    log("generator s16 ef16")
    g = lt.Graph.rmat(16, 16)
    rowptr, col = oracle.rmat_graph(16, 16)
    assert np.array_equal(g.rowptr(), rowptr)
    assert np.array_equal(g.col(), col)


CONFIGS = [
    # (label, partition bytes, engine kwargs)
    ("one-partition", 32 << 20, {}),
    ("many-resident", 16 << 10, {}),
    ("capped", 16 << 10, {"num_device_partitions": 3}),
    ("capped-zc", 16 << 10, {"num_device_partitions": 3, "zerocopy_threshold": 2000}),
    ("all-zc", 16 << 10, {"num_device_partitions": 2, "zerocopy_threshold": 1 << 40}),
    ("fifo-nopipe", 16 << 10, {"num_device_partitions": 4, "policy": "fifo", "pipeline": 0, "zerocopy_threshold": 0}),
    ("waves", 16 << 10, {"num_device_partitions": 5, "num_device_walkers": 3000}),
    ("span-resident", 16 << 10, {"num_device_partitions": 6, "span": "resident"}),
    ("device-sched", 16 << 10, {"num_device_partitions": 3, "device_schedule": 1}),
    ("device-sched-zc", 16 << 10, {"num_device_partitions": 4, "device_schedule": 1, "zerocopy_threshold": 2000,
                                   "span": "resident"}),
    ("small-slices", 16 << 10, {"num_device_partitions": 4, "slice_chunks": 1, "num_device_walkers": 9000}),
    # zero-copy batches: several sub-threshold queues walked over PCIe in one launch
    ("zc-batch", 16 << 10, {"num_device_partitions": 3, "span": "resident", "zerocopy_threshold": 4000,
                            "zc_batch": 16}),
    ("zc-batch-all", 16 << 10, {"num_device_partitions": 2, "span": "resident", "zerocopy_threshold": 1 << 40,
                                "zc_batch": 5, "slice_chunks": 2}),
    ("learned-fallback-batch", 16 << 10, {"num_device_partitions": 3, "span": "resident", "policy": "learned",
                                          "zerocopy_threshold": 3000}),
    # more walkers than the device holds: queues spill to host memory and come back
    ("spill", 16 << 10, {"num_device_partitions": 4, "num_device_walkers": 5000}),
    ("spill-resident-zc", 16 << 10, {"num_device_partitions": 3, "span": "resident", "zerocopy_threshold": 3000,
                                     "num_device_walkers": 7000, "zc_batch": 16}),
    ("spill-tiny", 16 << 10, {"num_device_partitions": 2, "num_device_walkers": 1000, "slice_chunks": 1}),
    ("waves-device-sched", 16 << 10, {"num_device_partitions": 3, "device_schedule": 1, "num_device_walkers": 3000}),
    # rowptr of the whole graph in HBM: zero-copied moves read only col over PCIe
    ("rp-resident-zc", 16 << 10, {"num_device_partitions": 6, "zerocopy_threshold": 4000, "rowptr_resident": 1}),
    ("rp-resident-batch", 16 << 10, {"num_device_partitions": 6, "span": "resident", "zerocopy_threshold": 1 << 40,
                                     "zc_batch": 16, "rowptr_resident": 1}),
    ("rp-resident-device", 16 << 10, {"num_device_partitions": 6, "device_schedule": 1, "zerocopy_threshold": 3000,
                                      "span": "resident", "rowptr_resident": 1}),
    # partitions copied packed and unpacked in HBM
    ("packed-capped", 16 << 10, {"num_device_partitions": 3, "pack_partitions": 1}),
    ("packed-learned-zc", 16 << 10, {"num_device_partitions": 3, "span": "resident", "policy": "learned",
                                     "zerocopy_threshold": 3000, "pack_partitions": 1}),
    ("packed-rp-resident", 16 << 10, {"num_device_partitions": 6, "zerocopy_threshold": 4000, "rowptr_resident": 1,
                                      "pack_partitions": 1}),
    ("packed-spill", 16 << 10, {"num_device_partitions": 4, "num_device_walkers": 5000, "pack_partitions": 1}),
    # partitions kept packed in HBM as 16-edge blocks, decoded per read
    ("blocks-capped", 16 << 10, {"num_device_partitions": 3, "pack_partitions": 2}),
    ("blocks-learned-zc", 16 << 10, {"num_device_partitions": 3, "span": "resident", "policy": "learned",
                                     "zerocopy_threshold": 3000, "pack_partitions": 2}),
    ("blocks-batch", 16 << 10, {"num_device_partitions": 4, "span": "resident", "zerocopy_threshold": 3000,
                                "zc_batch": 16, "pack_partitions": 2}),
    ("blocks-rp-resident", 16 << 10, {"num_device_partitions": 6, "zerocopy_threshold": 4000, "rowptr_resident": 1,
                                      "pack_partitions": 2}),
    ("blocks-spill", 16 << 10, {"num_device_partitions": 4, "num_device_walkers": 5000, "pack_partitions": 2}),
    # the device-scheduled loop copying block images into block slots
    ("blocks-device-sched", 16 << 10, {"num_device_partitions": 4, "device_schedule": 1, "pack_partitions": 2}),
    ("blocks-device-sched-zc", 16 << 10, {"num_device_partitions": 4, "device_schedule": 1, "pack_partitions": 2,
                                          "zerocopy_threshold": 2000, "span": "resident"}),
    # several resident queues walked by one launch (span resident)
    ("resident-batch", 16 << 10, {"num_device_partitions": 6, "span": "resident", "resident_batch": 16}),
    ("resident-batch-blocks-zc", 16 << 10, {"num_device_partitions": 4, "span": "resident", "resident_batch": 4,
                                            "zc_batch": 16, "zerocopy_threshold": 3000, "pack_partitions": 2}),
    ("resident-batch-all-cached", 16 << 10, {"span": "resident", "resident_batch": 16, "walkers_per_thread": 2}),
    # the prefetching loop without a network: largest-queue targets, plain victims
    ("prefetch-rule-blocks", 16 << 10, {"num_device_partitions": 4, "span": "resident", "policy": "learned",
                                        "prefetch_plain": 1, "zerocopy_threshold": 3000, "pack_partitions": 2}),
    ("prefetch-rule-depth3", 16 << 10, {"num_device_partitions": 6, "span": "resident", "policy": "learned",
                                        "prefetch_plain": 1, "prefetch_depth": 3, "resident_batch": 4}),
    # fixed walkers per walk-kernel lane, host and device loops
    ("wpt4-capped", 16 << 10, {"num_device_partitions": 3, "walkers_per_thread": 4}),
    ("wpt2-blocks-batch", 16 << 10, {"num_device_partitions": 4, "span": "resident", "zerocopy_threshold": 3000,
                                     "zc_batch": 16, "pack_partitions": 2, "walkers_per_thread": 2}),
    ("wpt4-device-sched-blocks", 16 << 10, {"num_device_partitions": 4, "device_schedule": 1, "pack_partitions": 2,
                                            "walkers_per_thread": 4}),
    ("wpt2-device-sched-zc", 16 << 10, {"num_device_partitions": 4, "device_schedule": 1, "zerocopy_threshold": 2000,
                                        "span": "resident", "walkers_per_thread": 2}),
    # the reference's figure-12 appends: each walker stored from its own lane
    # (NO_RANDOM_ACCESS_REDUCE), one queue reservation per walker (ONE_LEVEL_QUEUE)
    ("append-no-staging", 16 << 10, {"num_device_partitions": 3, "append_mode": 1}),
    ("append-one-level", 16 << 10, {"num_device_partitions": 3, "append_mode": 2}),
    ("append-one-level-batches", 16 << 10, {"num_device_partitions": 4, "span": "resident", "zc_batch": 16,
                                            "resident_batch": 16, "zerocopy_threshold": 3000, "pack_partitions": 2,
                                            "append_mode": 2, "walkers_per_thread": 4}),
    ("append-no-staging-device", 16 << 10, {"num_device_partitions": 4, "device_schedule": 1, "append_mode": 1,
                                            "walkers_per_thread": 2}),
    # partitions sized for block slots (lt_partition_build_blocks)
    ("balanced-blocks-capped", ("blocks", 12 << 10), {"num_device_partitions": 3, "pack_partitions": 2}),
    ("balanced-raw-slots", ("blocks", 12 << 10), {"num_device_partitions": 3, "zerocopy_threshold": 1500}),
    ("balanced-prefetch-rule", ("blocks", 12 << 10), {"num_device_partitions": 3, "span": "resident",
                                                      "policy": "learned", "prefetch_plain": 1, "prefetch_depth": 2,
                                                      "zerocopy_threshold": 3000, "pack_partitions": 2,
                                                      "zc_batch": 16, "resident_batch": 16, "walkers_per_thread": 1,
                                                      "rowptr_resident": 1}),
    ("balanced-device-sched", ("blocks", 12 << 10), {"num_device_partitions": 3, "device_schedule": 1,
                                                     "pack_partitions": 2}),
]


def table(g, pbytes):
    """The partition table of a case: the reference's split of pbytes bytes,
    or ("blocks", n): partitions sized for n-byte block slots."""
    if isinstance(pbytes, tuple):
        return lt.Partition.build_blocks(g, pbytes[1])
    return lt.Partition.build(g, pbytes)


@pytest.mark.parametrize("label,pbytes,kw", CONFIGS, ids=[c[0] for c in CONFIGS])
def test_pagerank_bit_exact(g12, label, pbytes, kw):
    log(f"pagerank {label}")
    part = table(g12, pbytes)
    eng = lt.Engine(g12, part, **kw)
    w, l = 2 * g12.num_node, 10
    counts, ends = eng.pagerank(w, l, endpoints=True)
    rc, re, rst = oracle.pagerank(g12.rowptr(), g12.col(), w, l, endpoints=True)
    st = eng.stats()
    assert counts.sum() == w * l
    assert np.array_equal(counts, rc)
    assert np.array_equal(ends, re)
    assert st["moves"] == rst[0] and st["teleports"] == rst[1] and st["visits"] == rst[2]


@pytest.mark.parametrize("label,pbytes,kw", CONFIGS, ids=[c[0] for c in CONFIGS])
def test_ppr_bit_exact(g12, label, pbytes, kw):
    log(f"ppr {label}")
    part = table(g12, pbytes)
    eng = lt.Engine(g12, part, **kw)
    w = 3 * g12.num_node
    for src in (0, g12.num_node - 1):
        counts, ends = eng.ppr(w, source=src, prob=0.15, endpoints=True)
        rc, re, rst = oracle.ppr(g12.rowptr(), g12.col(), w, source=src, prob=0.15, endpoints=True)
        assert counts.sum() == w
        assert np.array_equal(counts, rc)
        assert np.array_equal(ends, re)
        assert eng.stats()["moves"] == rst[0]


@pytest.mark.parametrize("label,pbytes,kw", CONFIGS, ids=[c[0] for c in CONFIGS])
def test_genericwalk_bit_exact(g12, label, pbytes, kw):
    log(f"genericwalk {label}")
    part = table(g12, pbytes)
    eng = lt.Engine(g12, part, **kw)
    w, l = 2 * g12.num_node, 10
    ends, moves = eng.genericwalk(w, l, runs=2, moves=True)
    re, rm, rst = oracle.genericwalk(g12.rowptr(), g12.col(), w, l, runs=2)
    assert np.array_equal(ends, re)
    assert np.array_equal(moves, rm)


@pytest.mark.parametrize("label,pbytes,kw", [c for c in CONFIGS if c[0] in (
    "one-partition", "capped", "all-zc", "span-resident", "device-sched-zc", "small-slices", "zc-batch",
    "packed-capped", "packed-learned-zc")],
    ids=lambda x: x if isinstance(x, str) else "")
def test_genericwalk_paths_bit_exact(g12, label, pbytes, kw):
    """Whole paths (layout of src/walk/NodeSampler.h:36-49) against the oracle."""
    log(f"genericwalk paths {label}")
    part = table(g12, pbytes)
    eng = lt.Engine(g12, part, **kw)
    w, l = 3 * g12.num_node, 9
    paths = eng.genericwalk_paths(w, l, runs=2)
    assert np.array_equal(paths, oracle.genericwalk_paths(g12.rowptr(), g12.col(), w, l, runs=2))


def test_genericwalk_paths_with_dead_ends():
    log("genericwalk paths dangling")
    g = _dangling_graph()
    part = lt.Partition.build(g, 2048)
    eng = lt.Engine(g, part, num_device_partitions=3, zerocopy_threshold=100)
    # unsorted neighbour lists: packed images hold blocks stored absolute
    packed = lt.Engine(g, part, num_device_partitions=3, zerocopy_threshold=0, pack_partitions=1)
    blocks = lt.Engine(g, part, num_device_partitions=3, zerocopy_threshold=0, pack_partitions=2)
    for l in (0, 1, 7):
        paths = eng.genericwalk_paths(4000, l)
        ref = oracle.genericwalk_paths(g.rowptr(), g.col(), 4000, l)
        assert np.array_equal(paths, ref)
        assert np.array_equal(packed.genericwalk_paths(4000, l), ref)
        assert np.array_equal(blocks.genericwalk_paths(4000, l), ref)
    assert (ref == 0xFFFFFFFF).any()
    assert packed.stats()["packed_image_bytes"] > 0
    assert blocks.stats()["packed_image_bytes"] > 0


def _dangling_graph():
    rng = np.random.default_rng(5)
    V = 3000
    deg = rng.integers(0, 6, V)
    deg[rng.random(V) < 0.2] = 0
    rowptr = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint64)
    col = rng.integers(0, V, int(rowptr[-1])).astype(np.uint32)
    return lt.Graph.from_csr(rowptr, col)


@pytest.mark.parametrize("kw", [{}, {"num_device_partitions": 2, "zerocopy_threshold": 100}])
def test_dangling_vertices(kw):
    log(f"dangling {kw}")
    g = _dangling_graph()
    part = lt.Partition.build(g, 2048)
    eng = lt.Engine(g, part, **kw)
    w = 5000
    counts, ends = eng.ppr(w, source=1, endpoints=True)
    rc, re, _ = oracle.ppr(g.rowptr(), g.col(), w, source=1, endpoints=True)
    assert np.array_equal(counts, rc) and np.array_equal(ends, re)
    ends, moves = eng.genericwalk(w, 7, moves=True)
    re, rm, _ = oracle.genericwalk(g.rowptr(), g.col(), w, 7)
    assert np.array_equal(ends, re) and np.array_equal(moves, rm)
    counts, ends = eng.pagerank(w, 7, endpoints=True)
    rc, re, _ = oracle.pagerank(g.rowptr(), g.col(), w, 7, endpoints=True)
    assert np.array_equal(counts, rc) and np.array_equal(ends, re)


@pytest.mark.parametrize("kw", [{}, {"device_schedule": 1}, {"device_schedule": 1, "span": "resident"},
                                {"span": "resident", "zerocopy_threshold": 5000, "zc_batch": 16},
                                # the C4 autotuner's usual pick: packed copies, resident span, batches
                                {"span": "resident", "zerocopy_threshold": 500, "zc_batch": 16,
                                 "pack_partitions": 1},
                                {"pack_partitions": 1, "zerocopy_threshold": 0, "slice_chunks": 1},
                                {"span": "resident", "zerocopy_threshold": 500, "zc_batch": 16,
                                 "pack_partitions": 2}])
def test_ppr_batch_bit_exact(g12, kw):
    log(f"ppr batch {kw}")
    part = lt.Partition.build(g12, 16 << 10)
    eng = lt.Engine(g12, part, num_device_partitions=4, **kw)
    rng = np.random.default_rng(9)
    nq = 64
    sources = rng.integers(0, g12.num_node, nq).astype(np.uint32)
    qc = np.exp(rng.uniform(np.log(10), np.log(3000), nq)).astype(np.uint64)
    ends = eng.ppr_batch(sources, qc, prob=0.15)
    re, _ = oracle.ppr_batch(g12.rowptr(), g12.col(), sources, qc, prob=0.15)
    assert np.array_equal(ends, re)


def test_autotuner_picks_a_trial():
    """The batched-PPR autotuner tries every configuration on a sample, times
    the three fastest and the fastest of each copy mode again on a larger
    sample, then the fastest of those with fixed walkers per lane, and
    returns the fastest of the last two stages; every trial walks the same
    queries."""
    from lighttraffic_amd.tune import tune_ppr_batch
    log("autotuner")
    g = lt.Graph.rmat(14, 8)
    src, cnt = g.ppr_queries(256, lo=100, hi=10000, seed=5)
    best, trials = tune_ppr_batch(g, src, cnt, cache_bytes=48 << 20, sample_every=4, sizes_mb=(16,))
    # host scheduling: (span partition x 2 slices + span resident x 2 slices x 2 batch
    # sizes) x raw, gap-coded or block copies; device scheduling: 2 spans (whole
    # queues) x raw or block copies; then block copies on the split for block
    # slots: 6 host and 2 device configurations
    first = [t for t in trials if t["stage"] == 1]
    refined = [t for t in trials if t["stage"] == 2]
    assert len(first) == 3 * (2 + 4) + 2 * 2 + (6 + 2)
    assert {t["pack_partitions"] for t in first if t["split"]} == {2}
    assert sum(t["split"] for t in first) == 8
    assert {(t["span"], t["pack_partitions"]) for t in first if t["device_schedule"]} == {
        ("partition", 0), ("partition", 2), ("resident", 0), ("resident", 2)}
    keys = ("partition_mb", "span", "device_schedule", "slice_chunks", "zc_batch", "pack_partitions", "split")
    want = {tuple(t[k] for k in keys) for t in sorted(first, key=lambda t: t["running_ms"])[:3]}
    for pk, split in ((0, 0), (1, 0), (2, 0), (2, 1)):
        want.add(tuple(min((t for t in first if t["pack_partitions"] == pk and t["split"] == split),
                           key=lambda t: t["running_ms"])[k] for k in keys))
    assert sorted(tuple(t[k] for k in keys) for t in refined) == sorted(want)
    assert {t["pack_partitions"] for t in refined} == {0, 1, 2}
    # stage 3: stage 2's pick with 1, 2 and 4 walkers per lane (the device
    # loop's default is 1, so it is not timed twice there)
    third = [t for t in trials if t["stage"] == 3]
    pick2 = min(refined, key=lambda t: t["running_ms"])
    assert pick2["walkers_per_thread"] == 0
    assert sorted(t["walkers_per_thread"] for t in third) == ([2, 4] if pick2["device_schedule"] else [1, 2, 4])
    assert all(tuple(t[k] for k in keys) == tuple(pick2[k] for k in keys) for t in third)
    assert best in refined + third and best["running_ms"] == min(t["running_ms"] for t in refined + third)
    assert all(t["cache_slots"] >= t["slots"] and t["iterations"] > 0 for t in trials)
    assert {t["zc_batch"] for t in first} == {1, 16}
    assert {t["pack_partitions"] for t in first} == {0, 1, 2}


def test_resident_batches_walk_several_queues():
    """With every partition cached (a second job) and span resident, the
    first walks find the walkers created in every partition's queue:
    resident_batch 16 walks
    up to 16 of those queues (those of at least 4096 walkers) per launch, so
    there are fewer launches, with counts identical to the oracle and to
    resident_batch 1."""
    log("resident batches")
    g = lt.Graph.rmat(16, 16)
    part = lt.Partition.build(g, 1 << 20)
    w = 2 * g.num_node
    assert part.count >= 3 and w // part.count >= 2 * 4096
    rc, re, _ = oracle.pagerank(g.rowptr(), g.col(), w, 10, endpoints=True)
    launches = {}
    for rb in (1, 16):
        eng = lt.Engine(g, part, span="resident", resident_batch=rb)
        eng.pagerank(w, 10)  # loads every partition; the second job starts with all of them cached
        counts, ends = eng.pagerank(w, 10, endpoints=True)
        st = eng.stats()
        eng.close()
        assert np.array_equal(counts, rc) and np.array_equal(ends, re), rb
        launches[rb] = st["kernel_launches"]
        log(f"  resident_batch {rb}: {st['kernel_launches']} launches, {st['resident_batched']} queues batched")
        assert (st["resident_batched"] > 0) == (rb > 1)
    assert launches[16] < launches[1]


def test_run_arguments_are_checked(g12):
    """Bad run arguments are refused with LT_E_ARG before any walk, and the
    engine runs normally afterwards."""
    log("run argument checks")
    part = lt.Partition.build(g12, 16 << 10)
    eng = lt.Engine(g12, part, num_device_partitions=3)
    V = g12.num_node
    bad = [lambda: eng.ppr(100, source=V), lambda: eng.ppr(100, prob=0.0), lambda: eng.ppr(100, prob=1.5),
           lambda: eng.ppr(100, prob=float("nan")), lambda: eng.pagerank(0, 10), lambda: eng.pagerank(4, 10, runs=0),
           lambda: eng.ppr_batch([0, 1], [0, 0]), lambda: eng.ppr_batch([0, V], [5, 5]),
           lambda: eng.ppr_batch([1], [10], prob=2.0)]
    for k, f in enumerate(bad):
        with pytest.raises(lt.LTError) as e:
            f()
        assert e.value.status == 6, k  # LT_E_ARG
    counts = eng.ppr(1000, source=7, prob=1.0)  # the largest probability: walks of one move
    rc, _, _ = oracle.ppr(g12.rowptr(), g12.col(), 1000, source=7, prob=1.0)
    assert np.array_equal(counts, rc) and int(counts.sum()) == 1000
    eng.close()


def test_schedule_independence(g12):
    """Identical results whatever the cache size, policy and span."""
    log("schedule independence")
    w = 2 * g12.num_node
    results = []
    for pbytes, kw in [(64 << 10, {}), (8 << 10, {"num_device_partitions": 2}),
                       (8 << 10, {"num_device_partitions": 7, "policy": "fifo", "span": "resident"})]:
        part = table(g12, pbytes)
        eng = lt.Engine(g12, part, **kw)
        results.append(eng.pagerank(w, 12, runs=2))
    for r in results[1:]:
        assert np.array_equal(r, results[0])


def test_c1_pagerank_s16_one_partition():
    """R-MAT scale 16 (ef 16), one 32 MB partition, w = 2|V|, l = 10."""
    log("C1 s16 pagerank")
    g = lt.Graph.rmat(16, 16)
    part = lt.Partition.build(g, 32 << 20)
    assert part.count == 1
    eng = lt.Engine(g, part)
    w, l = 2 * g.num_node, 10
    counts, ends = eng.pagerank(w, l, endpoints=True)
    rc, re, _ = oracle.pagerank(g.rowptr(), g.col(), w, l, endpoints=True)
    assert counts.sum() == w * l
    assert np.array_equal(counts, rc) and np.array_equal(ends, re)


def test_s20_genericwalk_subsample():
    """Large graph, many 1 MB partitions, capped cache: sampled walker ids."""
    log("s20 genericwalk subsample")
    g = lt.Graph.rmat(20, 16)
    part = lt.Partition.build(g, 1 << 20)
    eng = lt.Engine(g, part, num_device_partitions=part.count // 2)
    w, l = 2 * g.num_node, 10
    ends, moves = eng.genericwalk(w, l, moves=True)
    rng = np.random.default_rng(1)
    for lo in rng.integers(0, w - 5000, 4):
        lo = int(lo)
        re, rm, _ = oracle.genericwalk(g.rowptr(), g.col(), w, l, id_begin=lo, id_end=lo + 5000)
        assert np.array_equal(ends[lo:lo + 5000], re) and np.array_equal(moves[lo:lo + 5000], rm)
    st = eng.stats()
    assert st["moves"] == int(moves.sum())


def test_generator_dedupe_seams(monkeypatch):
    """Deduplication in small pieces gives the same graph (pieces meet inside runs of equal keys)."""
    log("generator dedupe pieces")
    rowptr, col = oracle.rmat_graph(12, 16)
    for piece in ("1000", "4096", "7"):
        monkeypatch.setenv("LT_RMAT_DEDUPE_PIECE", piece)
        g = lt.Graph.rmat(12, 16)
        assert np.array_equal(g.rowptr(), rowptr) and np.array_equal(g.col(), col)


def _many_partition_graph():
    """60,000 vertices of degree 0-8: small partition sizes give thousands of partitions."""
    rng = np.random.default_rng(21)
    V = 60000
    deg = rng.integers(0, 9, V)
    rowptr = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint64)
    col = rng.integers(0, V, int(rowptr[-1])).astype(np.uint32)
    return lt.Graph.from_csr(rowptr, col)


@pytest.mark.parametrize("S", [600, 150])
@pytest.mark.parametrize("kw", [{"num_device_partitions": 600},
                                {"num_device_partitions": 600, "span": "resident", "zerocopy_threshold": 40,
                                 "zc_batch": 16},
                                {"num_device_partitions": 600, "device_schedule": 1},
                                # the reference's figure-12 appends on the hashed path
                                {"num_device_partitions": 600, "append_mode": 1},
                                {"num_device_partitions": 600, "span": "resident", "append_mode": 2}])
def test_many_partitions(S, kw):
    """About 2,400 and 12,000 partitions (above 1,024 the append stage numbers
    each pass's destinations through an LDS hash table; the reference's
    constant table holds 12,287, src/partition/PartitionStrategy.h:13) with
    a capped cache."""
    g = _many_partition_graph()
    part = lt.Partition.build(g, S)
    log(f"P={part.count} {kw}")
    assert part.count > 2000 if S == 600 else part.count > 11000
    eng = lt.Engine(g, part, **kw)
    w = 100000
    counts, ends = eng.pagerank(w, 8, endpoints=True)
    rc, re, _ = oracle.pagerank(g.rowptr(), g.col(), w, 8, endpoints=True)
    assert np.array_equal(counts, rc) and np.array_equal(ends, re)
    counts = eng.ppr(w, source=5)
    assert np.array_equal(counts, oracle.ppr(g.rowptr(), g.col(), w, source=5)[0])
    ends, _ = eng.genericwalk(w, 12)
    assert np.array_equal(ends, oracle.genericwalk(g.rowptr(), g.col(), w, 12)[0])


def test_c5_full_size():
    """The C5 workload at full size: R-MAT s26 ef27, 429 partitions of 32 MB,
    an 8 GB cache (256 slots), pagerank w = 2|V|, l = 10.  The reference
    scheduler's run matches the oracle's full u32 visit counts and every
    endpoint; the learned scheduler's run, with raw copies, gap-coded copies
    and blocks kept packed in HBM, gives the same visit counts (and
    endpoints), and so do the plain and learned runs on the graph split for
    32 MiB block slots."""
    log("C5 full size")
    g = lt.Graph.rmat(26, 27)
    part = lt.Partition.build(g, 32 << 20)
    w, l = 2 * g.num_node, 10
    plain = lt.Engine(g, part, num_device_partitions=256, record_trace=1)
    counts, ends = plain.pagerank(w, l, endpoints=True)
    assert part.count > 256 and int(counts.sum()) == w * l
    assert plain.stats()["graph_loading_calls"] > 0 and plain.stats()["zerocopy_calls"] > 0
    log("C5 oracle, all walkers")
    rc, re, _ = oracle.pagerank(g.rowptr(), g.col(), w, l, endpoints=True, threads=16)
    assert np.array_equal(ends, re)
    assert np.array_equal(counts, rc)
    sc = lt.Scorer(part, seed=1)
    sc.add_trace(plain, True, True)
    sc.fit(epochs=3)
    plain.close()
    learned = lt.Engine(g, part, num_device_partitions=256, span="resident", policy="learned")
    learned.set_scorer(sc)
    assert np.array_equal(learned.pagerank(w, l), counts)
    assert learned.stats()["prefetch_issued"] > 0
    learned.close()
    packed = lt.Engine(g, part, num_device_partitions=256, span="resident", policy="learned", pack_partitions=1)
    packed.set_scorer(sc)
    pc, pe = packed.pagerank(w, l, endpoints=True)
    assert np.array_equal(pc, counts) and np.array_equal(pe, ends)
    st = packed.stats()
    assert st["packed_image_bytes"] > 0 and st["packed_col_bytes"] > 1.5 * st["packed_image_bytes"]
    packed.close()
    # blocks kept packed in the slots of the 32 MB split
    blocks = lt.Engine(g, part, num_device_partitions=256, span="resident", policy="learned", pack_partitions=2)
    blocks.set_scorer(sc)
    assert np.array_equal(blocks.pagerank(w, l), counts)
    assert blocks.stats()["cache_slots"] > 256
    blocks.close()
    g.drop_packed()
    # the bench's headline configuration: the graph split for 32 MiB block
    # slots (the same 8 GB), blocks kept packed in the slots, largest-queue
    # prefetch targets at depth 2, the network's victims, threshold slot/128,
    # resident and zero-copy batches, 1 walker per lane
    pblk = lt.Partition.build_blocks(g, 32 << 20)
    slots = (8 << 30) // pblk.bytes
    bplain = lt.Engine(g, pblk, num_device_partitions=slots, span="resident", pack_partitions=2, record_trace=1,
                       zerocopy_threshold=(32 << 20) // 192)
    assert np.array_equal(bplain.pagerank(w, l), counts)
    sb = lt.Scorer(pblk, seed=1)
    sb.add_trace(bplain, True, 2)
    sb.fit(epochs=3)
    bplain.close()
    blocks = lt.Engine(g, pblk, num_device_partitions=slots, span="resident", policy="learned", pack_partitions=2,
                       prefetch_plain=2, prefetch_depth=2, zerocopy_threshold=(32 << 20) // 128,
                       walkers_per_thread=1)
    blocks.set_scorer(sb)
    bc, be = blocks.pagerank(w, l, endpoints=True)
    assert np.array_equal(bc, counts) and np.array_equal(be, ends)
    st = blocks.stats()
    log(f"  balanced split: P={pblk.count}, {st['cache_slots']} slots, {st['running_ms']:.1f} ms")
    assert st["cache_slots"] > 0.8 * pblk.count and st["packed_image_bytes"] > 0
    blocks.close()


def test_spill_moves_walkers_through_host_memory(g12):
    """Walker spill (src/Scheduler.h:53-82, 215-233): with 2,000 device walker
    slots for 8,192 walkers, queues go to host memory and come back; results
    equal the oracle's, and each reloaded walker adds 16 bytes of H2D."""
    log("spill stats")
    part = lt.Partition.build(g12, 16 << 10)
    eng = lt.Engine(g12, part, num_device_partitions=3, num_device_walkers=2000, zerocopy_threshold=0)
    w = 2 * g12.num_node
    counts, ends = eng.pagerank(w, 10, endpoints=True)
    rc, re, _ = oracle.pagerank(g12.rowptr(), g12.col(), w, 10, endpoints=True)
    assert np.array_equal(counts, rc) and np.array_equal(ends, re)
    st = eng.stats()
    assert st["walker_evicting_calls"] > 0 and st["walker_loading_calls"] > 0
    assert st["h2d_bytes_zerocopy"] > 0 and st["h2d_bytes_zerocopy"] % 16 == 0  # no zero-copy graph reads here
    # the same engine without a cap keeps everything on the device
    eng2 = lt.Engine(g12, part, num_device_partitions=3, zerocopy_threshold=0)
    assert np.array_equal(eng2.pagerank(w, 10), rc)
    assert eng2.stats()["walker_evicting_calls"] == 0 and eng2.stats()["h2d_bytes_zerocopy"] == 0


def test_spill_with_learned_scheduler(g12):
    log("spill learned")
    part = lt.Partition.build(g12, 16 << 10)
    tr = lt.Engine(g12, part, num_device_partitions=3, span="resident", record_trace=1)
    w = 2 * g12.num_node
    ref = tr.pagerank(w, 10)
    sc = lt.Scorer(part, seed=2)
    sc.add_trace(tr, True, True)
    sc.fit(epochs=3)
    eng = lt.Engine(g12, part, num_device_partitions=3, span="resident", policy="learned", num_device_walkers=3000)
    eng.set_scorer(sc)
    assert np.array_equal(eng.pagerank(w, 10), ref)
    assert eng.stats()["walker_evicting_calls"] > 0


def test_resident_rowptr_halves_zero_copy_reads(g12):
    """With rowptr in HBM a zero-copied move reads one 64-byte line (its col
    entry) instead of two; the rowptr upload is charged once, to the first run."""
    log("resident rowptr bytes")
    part = lt.Partition.build(g12, 16 << 10)
    w = 2 * g12.num_node
    kw = dict(num_device_partitions=6, zerocopy_threshold=1 << 40)
    a = lt.Engine(g12, part, rowptr_resident=-1, **kw)
    b = lt.Engine(g12, part, rowptr_resident=1, **kw)
    ca, cb = a.pagerank(w, 10), b.pagerank(w, 10)
    assert np.array_equal(ca, cb)
    sa, sb = a.stats(), b.stats()
    assert sa["zerocopy_moves"] == sb["zerocopy_moves"] > 0
    assert sa["h2d_bytes_zerocopy"] == 128 * sa["zerocopy_moves"]
    assert sb["h2d_bytes_zerocopy"] == 64 * sb["zerocopy_moves"]
    assert sb["h2d_bytes_explicit"] == 8 * (g12.num_node + 1)  # the upload; nothing else is copied
    b.pagerank(w, 10)
    assert b.stats()["h2d_bytes_explicit"] == 0
