"""Host-side tests of liblighttraffic.so (no GPU needed).

- the library loads and exports every function include/lighttraffic.h declares;
- bcsr and .part files: byte layout (src/GraphLoader.h:30-48,
  src/partition/PartitionStrategy.h:76-97, 325-337) and round trips;
- the greedy partitioner against the oracle and its invariants
  (src/partition/PartitionStrategy.h:189-235);
- the pure scheduler rules against the oracle (src/Scheduler.h:84-171, 368-388);
- error paths: missing/truncated files, oversize vertex, no device.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import lighttraffic_amd as lt
from oracle import ffi as oracle

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLDEN = os.path.join(HERE, "golden")
HEADER = os.path.join(ROOT, "include", "lighttraffic.h")


def _no_device():
    try:
        lt.device_info()
        return False
    except lt.LTError:
        return True


def test_every_declared_symbol_is_exported():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(lt_[a-z0-9_]+)\s*\(", text))
    assert len(names) > 30
    out = subprocess.run(["nm", "-D", "--defined-only", lt.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (lt_\w+)", out))
    assert names <= exported, sorted(names - exported)


def test_status_strings_and_defaults():
    L = lt.lib()
    for code in range(9):
        assert L.lt_status_str(code)
    c = lt.default_config()
    assert c.zerocopy_threshold == 2 ** 64 - 1 and c.pipeline == 1 and c.policy == 0


@pytest.mark.parametrize("name", ["hand", "rmat_s8_ef4", "rmat_s10_ef8"])
def test_bcsr_open_and_roundtrip(name, tmp_path):
    path = os.path.join(GOLDEN, name + ".bcsr")
    g = lt.Graph.open(path)
    rowptr, col = oracle.read_bcsr(path)
    assert g.num_node == len(rowptr) - 1 and g.num_edge == len(col)
    assert np.array_equal(g.rowptr(), rowptr) and np.array_equal(g.col(), col)
    out = tmp_path / "x.bcsr"
    g.save(str(out))
    assert out.read_bytes() == open(path, "rb").read()
    # header: u32 V, u64 E, no padding
    raw = out.read_bytes()
    assert len(raw) == 12 + 8 * (g.num_node + 1) + 4 * g.num_edge


def test_bcsr_errors(tmp_path):
    with pytest.raises(lt.LTError) as e:
        lt.Graph.open(str(tmp_path / "missing.bcsr"))
    assert e.value.status == 1
    raw = open(os.path.join(GOLDEN, "hand.bcsr"), "rb").read()
    (tmp_path / "short.bcsr").write_bytes(raw[:-3])
    with pytest.raises(lt.LTError) as e:
        lt.Graph.open(str(tmp_path / "short.bcsr"))
    assert e.value.status == 2
    bad = bytearray(raw)
    bad[-1] = 0xFF  # a column id >= V
    (tmp_path / "bad.bcsr").write_bytes(bytes(bad))
    with pytest.raises(lt.LTError) as e:
        lt.Graph.open(str(tmp_path / "bad.bcsr"))
    assert e.value.status == 2


def test_from_csr_refuses_what_the_kernels_cannot_read():
    """A rowptr that decreases, or whose interior offsets pass E, would make a
    walk kernel read far outside the column slice: from_csr refuses it as
    lt_graph_open does (LT_E_FORMAT)."""
    col = np.zeros(8, np.uint32)
    for rowptr in ([0, 5, 3, 8], [0, 100, 100, 8]):
        with pytest.raises(lt.LTError) as e:
            lt.Graph.from_csr(np.array(rowptr, np.uint64), col)
        assert e.value.status == 2, rowptr
    with pytest.raises(lt.LTError) as e:  # a column id >= V
        lt.Graph.from_csr(np.array([0, 1, 2, 8], np.uint64), np.array([0, 1, 2, 3, 0, 0, 0, 0], np.uint32))
    assert e.value.status == 2
    g = lt.Graph.from_csr(np.array([0, 3, 3, 8], np.uint64), col)  # an empty row is fine
    assert g.num_node == 3 and g.num_edge == 8


def _random_csr(rng, V, maxdeg):
    deg = rng.integers(0, maxdeg, V)
    rowptr = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint64)
    col = rng.integers(0, V, int(rowptr[-1])).astype(np.uint32)
    return rowptr, col


def test_partition_matches_oracle():
    rng = np.random.default_rng(11)
    for trial in range(25):
        V = int(rng.integers(1, 3000))
        rowptr, col = _random_csr(rng, V, 40)
        g = lt.Graph.from_csr(rowptr, col)
        S = int(rng.integers(8 * 2 + 4 * 40, 20000))
        p = lt.Partition.build(g, S)
        no, eo = oracle.partition(rowptr, S)
        assert np.array_equal(p.node_offsets(), no) and np.array_equal(p.edge_offsets(), eo)
        for v in rng.integers(0, V, 50):
            assert p.find(int(v)) == oracle.find_partition(no, int(v))


def test_partition_rmat_golden_and_counts():
    g = lt.Graph.open(os.path.join(GOLDEN, "rmat_s10_ef8.bcsr"))
    total = 8 * (g.num_node + 1) + 4 * g.num_edge
    for S in (4096, 10000, 1 << 16, 1 << 25):
        p = lt.Partition.build(g, S)
        no, eo = oracle.partition(g.rowptr(), S)
        assert np.array_equal(p.node_offsets(), no)
        # greedy count is at least ceil(total/S) (each part <= S, parts share no rows)
        assert p.count >= -(-total // S)
    assert lt.Partition.build(g, 1 << 25).count == 1


def test_partition_file_layout_and_roundtrip(tmp_path):
    src = os.path.join(GOLDEN, "rmat_s10_ef8.bcsr")
    gpath = str(tmp_path / "g.bcsr")
    open(gpath, "wb").write(open(src, "rb").read())
    g = lt.Graph.open(gpath)
    p = lt.Partition.build(g, 8192)
    p.save(gpath)
    raw = open(gpath + ".part", "rb").read()
    P = p.count
    assert len(raw) == 2 + 8 + 4 * (P + 1) + 8 * (P + 1)  # u16 P | u64 S | u32[P+1] | u64[P+1]
    assert int(np.frombuffer(raw[:2], np.uint16)[0]) == P
    assert int(np.frombuffer(raw[2:10], np.uint64)[0]) == 8192
    assert np.array_equal(np.frombuffer(raw[10:10 + 4 * (P + 1)], np.uint32), p.node_offsets())
    assert np.array_equal(np.frombuffer(raw[10 + 4 * (P + 1):], np.uint64), p.edge_offsets())
    q = lt.Partition.load(gpath)
    assert q.count == P and q.bytes == 8192
    assert np.array_equal(q.node_offsets(), p.node_offsets())
    assert np.array_equal(q.edge_offsets(), p.edge_offsets())


def test_more_than_65535_partitions_is_an_error():
    """Partition ids are u16 in .part files (src/partition/PartitionStrategy.h:76-97)
    and in the engine's append stage: a split into more is refused."""
    rng = np.random.default_rng(22)
    V = 200000
    deg = rng.integers(0, 5, V)
    rowptr = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint64)
    g = lt.Graph.from_csr(rowptr, rng.integers(0, V, int(rowptr[-1])).astype(np.uint32))
    assert lt.Partition.build(g, 1000).count <= 65535
    with pytest.raises(lt.LTError) as e:
        lt.Partition.build(g, 40)
    assert e.value.status == 6


def test_partition_errors(tmp_path):
    rowptr = np.array([0, 3, 1003, 1005], np.uint64)
    col = np.zeros(1005, np.uint32)
    g = lt.Graph.from_csr(rowptr, col)
    with pytest.raises(lt.LTError) as e:
        lt.Partition.build(g, 1000)
    assert e.value.status == 3  # LT_E_OVERSIZE_VERTEX
    with pytest.raises(lt.LTError) as e:
        lt.Partition.load(str(tmp_path / "nothing.bcsr"))
    assert e.value.status == 1


def test_scheduler_rules_match_oracle():
    rng = np.random.default_rng(2)
    for trial in range(300):
        P = int(rng.integers(1, 40))
        counts = rng.integers(0, 5, P) * rng.integers(0, 300000, P)
        if trial % 5 == 0:
            counts[:] = counts[0]  # ties
        resident = (rng.random(P) < 0.5).astype(np.uint8)
        prev = int(rng.integers(-1, P))
        for sync in (0, 1):
            assert lt.sched_greedy_next(counts, prev, sync) == oracle.greedy_next(counts, prev, sync)
        assert lt.sched_evict_min(counts, resident, prev) == oracle.evict_min(counts, resident, prev)
        page = int(rng.choice([1, 1000, 167936]))
        assert (lt.sched_computed_partition(counts, resident, prev, page) ==
                oracle.computed_partition(counts, resident, prev, page))


def test_scheduler_rules_known_cases():
    # greedy skips the previously computed partition unless synchronised
    assert lt.sched_greedy_next([5, 9, 9, 1], 1, 0) == 2
    assert lt.sched_greedy_next([5, 9, 9, 1], 1, 1) == 1
    assert lt.sched_greedy_next([0, 0, 0], -1, 1) == -1
    # eviction: fewest walkers among resident partitions except the last loaded
    assert lt.sched_evict_min([7, 1, 3, 0], [1, 1, 1, 0], 1) == 2
    # computed partition: a "full" queue (>= page) with the fewest walkers wins
    assert lt.sched_computed_partition([10, 500, 300, 50], [1, 1, 1, 1], 0, 200) == 2
    assert lt.sched_computed_partition([10, 100, 30, 50], [1, 1, 1, 1], 0, 200) == 1


@pytest.mark.skipif(not _no_device(), reason="a HIP device is visible")
def test_engine_without_device_fails_loudly():
    g = lt.Graph.open(os.path.join(GOLDEN, "rmat_s8_ef4.bcsr"))
    p = lt.Partition.build(g, 1 << 20)
    with pytest.raises(lt.LTError) as e:
        lt.Engine(g, p)
    assert e.value.status == 7  # LT_E_NODEVICE


def test_engine_rejects_unknown_pack_mode():
    # checked before any device call, so on every host
    g = lt.Graph.open(os.path.join(GOLDEN, "rmat_s8_ef4.bcsr"))
    p = lt.Partition.build(g, 1 << 20)
    for mode in (-1, 3):
        with pytest.raises(lt.LTError) as e:
            lt.Engine(g, p, pack_partitions=mode)
        assert e.value.status == 6 and "pack_partitions" in str(e.value)  # LT_E_ARG


def test_engine_rejects_unknown_walkers_per_thread():
    g = lt.Graph.open(os.path.join(GOLDEN, "rmat_s8_ef4.bcsr"))
    p = lt.Partition.build(g, 1 << 20)
    for wpt in (-1, 3, 8):
        with pytest.raises(lt.LTError) as e:
            lt.Engine(g, p, walkers_per_thread=wpt)
        assert e.value.status == 6 and "walkers_per_thread" in str(e.value)  # LT_E_ARG


@pytest.mark.parametrize("field,bad", [("prefetch_plain", -1), ("prefetch_plain", 3), ("evict_shortlist", -2),
                                       ("learned_next", -1), ("learned_next", 101), ("append_mode", 3)])
def test_engine_rejects_bad_learned_settings(field, bad):
    # checked before any device call, so on every host
    g = lt.Graph.open(os.path.join(GOLDEN, "rmat_s8_ef4.bcsr"))
    p = lt.Partition.build(g, 1 << 20)
    with pytest.raises(lt.LTError) as e:
        lt.Engine(g, p, policy="learned", **{field: bad})
    assert e.value.status == 6 and field in str(e.value)  # LT_E_ARG


def test_config_defaults_of_the_new_fields():
    c = lt.Config()
    lt.lib().lt_config_default(lt.ctypes.byref(c))
    assert (c.prefetch_plain, c.evict_shortlist, c.walkers_per_thread, c.resident_batch, c.learned_next,
            c.append_mode) == (0,) * 6
    assert (c.prefetch_depth, c.prefetch_stick_pct) == (2, 50)


@pytest.mark.skipif(not _no_device(), reason="a HIP device is visible")
def test_bandwidth_probes_without_device():
    import ctypes
    L = lt.lib()
    gbps = ctypes.c_double()
    assert L.lt_device_h2d_bandwidth(0, 1, ctypes.byref(gbps)) == 6  # LT_E_ARG
    assert L.lt_device_h2d_bandwidth(1 << 20, 1, ctypes.byref(gbps)) == 7  # LT_E_NODEVICE
    assert L.lt_device_copy_bandwidth(1 << 20, 1, ctypes.byref(gbps)) == 7


@pytest.mark.skipif(not _no_device(), reason="a HIP device is visible")
def test_cli_materialises_partitions_then_needs_a_device(tmp_path):
    gpath = str(tmp_path / "g.bcsr")
    open(gpath, "wb").write(open(os.path.join(GOLDEN, "rmat_s10_ef8.bcsr"), "rb").read())
    r = subprocess.run([os.path.join(lt.BIN_DIR, "pagerank"), "-f", gpath, "-m", "-s", "1"], capture_output=True,
                       text=True)
    assert r.returncode == 1
    assert "node with max degree:" in r.stdout and "num nodes:" in r.stdout
    assert "no HIP device" in r.stderr
    p = lt.Partition.load(gpath)
    q = lt.Partition.build(lt.Graph.open(gpath), 1 << 20)
    assert np.array_equal(p.node_offsets(), q.node_offsets())
    r = subprocess.run([os.path.join(lt.BIN_DIR, "ppr"), "-f", str(tmp_path / "none.bcsr")], capture_output=True,
                       text=True)
    assert r.returncode == 1 and "open graph" in r.stderr


def test_ppr_query_table():
    """Sources proportional to degree, counts log-uniform; restated with the
    oracle's Philox and checked statistically."""
    g = lt.Graph.open(os.path.join(GOLDEN, "rmat_s10_ef8.bcsr"))
    nq = 20000
    src, cnt = g.ppr_queries(nq, 1000, 1000000, seed=99)
    rowptr = g.rowptr()
    E = int(rowptr[-1])
    tag = 0x51525953
    for q in range(0, nq, 997):
        x = oracle.philox([q, 0, 0, tag], [99, 0])
        e = (((x[0] << 32) | x[1]) * E) >> 64
        assert src[q] == int(np.searchsorted(rowptr, e, side="right") - 1)
        u = x[2] / 2 ** 32
        assert cnt[q] == int(np.floor(np.exp(np.log(1000) + u * (np.log(1e6) - np.log(1000)))))
    assert cnt.min() >= 1000 and cnt.max() < 1000000
    # log-uniform: log10(count) is uniform on [3, 6)
    assert abs(np.log10(cnt.astype(float)).mean() - 4.5) < 0.02
    deg = np.diff(rowptr.astype(np.int64))
    hits = np.bincount(src, minlength=len(deg))
    exp = nq * deg / deg.sum()
    big = exp > 20
    assert np.all(np.abs(hits[big] - exp[big]) < 6 * np.sqrt(exp[big]))


def _min_loads_brute(seq, slots, start=()):
    """Fewest loads over every eviction choice (exhaustive search with memo)."""
    from functools import lru_cache

    @lru_cache(maxsize=None)
    def go(i, cache):
        if i == len(seq):
            return 0
        p = seq[i]
        if p in cache:
            return go(i + 1, cache)
        if len(cache) < slots:
            return 1 + go(i + 1, frozenset(cache | {p}))
        return 1 + min(go(i + 1, frozenset((cache - {v}) | {p})) for v in cache)

    return go(0, frozenset(start))


def _lru_loads(seq, slots):
    cache, loads = [], 0
    for p in seq:
        if p in cache:
            cache.remove(p)
        else:
            loads += 1
            if len(cache) == slots:
                cache.pop(0)
        cache.append(p)
    return loads


def test_belady_hand_computed():
    # 3 slots: 0 1 2 load; 3 evicts 2 (next use farthest); 2 evicts 0 (0 and 1
    # never used again, lowest id first); 3 hits
    loads, nbytes, victims = lt.sched_belady([0, 1, 2, 0, 1, 3, 0, 1, 2, 3], 4, 3, part_bytes=[10, 20, 30, 40])
    assert loads == 5 and nbytes == 10 + 20 + 30 + 40 + 30
    assert list(victims) == [-1, -1, -1, -1, -1, 2, -1, -1, 0, -1]
    # a warm start: partitions 0 and 2 resident, 2 slots
    loads, _, victims = lt.sched_belady([2, 1, 0, 1], 3, 2, init_resident=[1, 0, 1])
    assert loads == 1 and list(victims) == [-1, 2, -1, -1]
    with pytest.raises(lt.LTError) as e:
        lt.sched_belady([0, 5], 4, 2)
    assert e.value.status == 6


def test_belady_is_the_minimum():
    """Belady's loads equal the exhaustive minimum on small random sequences
    and never exceed LRU's."""
    rng = np.random.default_rng(7)
    for trial in range(300):
        P = int(rng.integers(2, 6))
        slots = int(rng.integers(1, P))
        seq = [int(x) for x in rng.integers(0, P, int(rng.integers(1, 13)))]
        start = tuple(sorted(set(int(x) for x in rng.integers(0, P, int(rng.integers(0, slots + 1))))))
        init = np.zeros(P, np.uint8)
        init[list(start)] = 1
        loads, _, victims = lt.sched_belady(seq, P, slots, init_resident=init)
        assert loads == _min_loads_brute(seq, slots, start), (seq, slots, start)
        if not start:
            assert loads <= _lru_loads(seq, slots)
        assert loads == sum(1 for i, p in enumerate(seq) if victims[i] >= 0) + min(
            slots - len(start), len(set(seq) - set(start)))  # every load past the free slots evicts
