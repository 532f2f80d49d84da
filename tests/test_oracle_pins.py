"""Pins of the CPU oracle against facts it does not compute itself.

- Philox4x32-10: Random123 known-answer vectors and vectors from ROCm's
  open-source rocrand header (tests/golden/philox_*.json).
- edge list -> bcsr: byte-exact against the reference's own converter
  (tools/convert_from_txt.cc, built into oracle/_ref) on committed fixtures.
- walk semantics: exact expectations from FP64 power iteration of the
  transition matrices (app/pagerank.cu:26-44, app/ppr.cu:22-42,
  app/genericwalk.cu:22-32), closed forms on regular graphs, and the
  reference's own sum invariants (app/pagerank.cu:94, app/ppr.cu:82).
"""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import ffi as oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
ROOT = os.path.dirname(HERE)


def _vectors(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)["vectors"]


@pytest.mark.parametrize("name", ["philox_kat.json", "philox_rocrand.json"])
def test_oracle_philox_known_answers(name):
    for v in _vectors(name):
        ctr = [int(x, 16) for x in v["ctr"]]
        key = [int(x, 16) for x in v["key"]]
        assert oracle.philox(ctr, key) == [int(x, 16) for x in v["out"]]


def test_engine_rng_header_known_answers(tmp_path):
    """The engine's lt_rng.h (host pass of the same code the kernels use)."""
    src = tmp_path / "kat.cpp"
    vecs = _vectors("philox_kat.json") + _vectors("philox_rocrand.json")
    lines = ['#include <cstdio>', '#include "lt_rng.h"', "int main() {"]
    for v in vecs:
        c = ", ".join("0x" + x + "u" for x in v["ctr"])
        k = v["key"]
        lines.append(f"  {{ lt::u32x4 r = lt::philox4x32_10(lt::u32x4{{{c}}}, 0x{k[0]}u, 0x{k[1]}u);"
                     ' printf("%08x %08x %08x %08x\\n", r.x, r.y, r.z, r.w); }')
    # bounded() and prob_threshold() closed forms
    lines.append('  printf("%llu\\n", (unsigned long long)lt::bounded(0xFFFFFFFFu, 0xFFFFFFFFu, 1000003ull));')
    lines.append('  printf("%llu\\n", (unsigned long long)lt::bounded(0x80000000u, 0u, 10ull));')
    lines.append('  printf("%u %u %u\\n", lt::prob_threshold(0.15), lt::prob_threshold(0.0), lt::prob_threshold(1.0));')
    lines.append("  return 0; }")
    src.write_text("\n".join(lines))
    exe = tmp_path / "kat"
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "lighttraffic_amd", "csrc"), str(src), "-o",
                    str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    for v, line in zip(vecs, out):
        assert line.split() == v["out"]
    tail = out[len(vecs):]
    r = (0xFFFFFFFF << 32) | 0xFFFFFFFF
    assert int(tail[0]) == (r * 1000003) >> 64
    assert int(tail[1]) == 5
    assert tail[2].split() == [str(int(0.15 * 2 ** 32)), "0", str(2 ** 32 - 1)]


def test_thresholds():
    assert oracle.L.ltor_threshold(0.15) == int(0.15 * 2 ** 32)
    # pagerank teleports when curand_uniform_double > 0.85f (app/pagerank.cu:6,35)
    assert oracle.L.ltor_teleport_threshold() == int((1.0 - float(np.float32(0.85))) * 2 ** 32)


@pytest.mark.parametrize("name", ["hand", "rmat_s8_ef4", "rmat_s10_ef8"])
def test_convert_matches_reference_converter(name):
    """ltor_convert == tools/convert_from_txt.cc on the same text edge list."""
    src, dst = [], []
    with open(os.path.join(GOLDEN, name + ".el")) as f:
        for line in f:
            if not line or not line[0].isdigit():
                continue
            a, b = line.split()[:2]
            src.append(int(a))
            dst.append(int(b))
    rowptr, col = oracle.convert(np.array(src, np.uint64), np.array(dst, np.uint64))
    with open(os.path.join(GOLDEN, name + ".bcsr"), "rb") as f:
        assert oracle.bcsr_bytes(rowptr, col) == f.read()


@pytest.mark.parametrize("name,scale,ef,seed", [("rmat_s8_ef4", 8, 4, 0x5EED0008), ("rmat_s10_ef8", 10, 8, 0x5EED000A)])
def test_rmat_sampler_fixture(name, scale, ef, seed):
    src, dst = oracle.rmat_edges(scale, ef, seed=seed)
    with open(os.path.join(GOLDEN, name + ".el")) as f:
        rows = [tuple(map(int, line.split())) for line in f if line.strip()]
    assert rows == list(zip(src.tolist(), dst.tolist()))


def test_rmat_quadrant_frequencies():
    """Every level picks (src bit, dst bit) with probabilities a, b, c, d."""
    a, b, c = 0.57, 0.19, 0.19
    scale = 12
    src, dst = oracle.rmat_edges(scale, 64, a=a, b=b, c=c, seed=7)
    n = len(src)
    for bit in (scale - 1, scale // 2, 0):
        sb = (src >> bit) & 1
        db = (dst >> bit) & 1
        for (x, y), p in {(0, 0): a, (0, 1): b, (1, 0): c, (1, 1): 1 - a - b - c}.items():
            k = int(np.sum((sb == x) & (db == y)))
            assert abs(k - n * p) < 6 * np.sqrt(n * p * (1 - p)), (bit, x, y)


# ---- walk semantics ---------------------------------------------------------

def _dense_P(rowptr, col):
    V = len(rowptr) - 1
    P = np.zeros((V, V))
    for v in range(V):
        b, e = int(rowptr[v]), int(rowptr[v + 1])
        for x in col[b:e]:
            P[v, x] += 1.0 / (e - b)
    return P


def _small_graph():
    rowptr, col = oracle.rmat_graph(7, 4, seed=11)
    return rowptr, col


TELE = oracle.L.ltor_teleport_threshold() / 2 ** 32


def test_pagerank_matches_power_iteration():
    rowptr, col = _small_graph()
    V = len(rowptr) - 1
    w, l = 400 * V, 6
    counts, _, st = oracle.pagerank(rowptr, col, w, l)
    assert counts.sum() == w * l  # app/pagerank.cu:94
    P = _dense_P(rowptr, col)
    Pa = (1 - TELE) * P + TELE / V
    x = np.full(V, 1.0 / V)
    E = np.zeros(V)
    for _ in range(l):
        E += x
        x = x @ Pa
    E *= w
    dev = np.abs(counts - E)
    assert np.all(dev <= 6 * np.sqrt(l * E) + 1), dev.max()
    assert np.abs(counts - E).sum() / (w * l) < 0.02
    assert abs(st[1] / (w * (l - 1)) - TELE) < 0.005  # teleport rate


def test_pagerank_regular_graph_closed_form():
    """2-regular cycle: every vertex expects w*l/V visits."""
    V = 64
    rowptr = np.arange(0, 2 * V + 1, 2, dtype=np.uint64)
    col = np.array([[(v - 1) % V, (v + 1) % V] for v in range(V)], np.uint32).ravel()
    w, l = 2000 * V, 5
    counts, _, _ = oracle.pagerank(rowptr, col, w, l)
    exp = w * l / V
    assert np.all(np.abs(counts - exp) <= 6 * np.sqrt(l * exp))


def _ppr_expectation(rowptr, col, s, p, steps=200):
    V = len(rowptr) - 1
    P = _dense_P(rowptr, col)
    dangling = (rowptr[1:] - rowptr[:-1]) == 0
    P[dangling, dangling] = 0
    for v in np.nonzero(dangling)[0]:
        P[v, v] = 1.0  # absorbing
    x = np.zeros(V)
    x[s] = 1.0
    r = np.zeros(V)
    for m in range(1, steps):
        x = x @ P
        r += p * (1 - p) ** (m - 1) * x
    return r


def test_ppr_matches_closed_form():
    rowptr, col = _small_graph()
    w = 200000
    p = oracle.L.ltor_threshold(float(np.float32(0.15))) / 2 ** 32
    counts, ends, st = oracle.ppr(rowptr, col, w, source=3, prob=0.15, endpoints=True)
    assert counts.sum() == w  # app/ppr.cu:82
    assert np.array_equal(np.bincount(ends, minlength=len(counts)), counts)
    r = _ppr_expectation(rowptr, col, 3, p) * w
    assert np.all(np.abs(counts - r) <= 6 * np.sqrt(r) + 1)
    # mean number of moves = 1/p (geometric length 1+F)
    assert abs(st[0] / w - 1 / p) < 0.05


def _dangling_graph():
    # 0->1->2->3 (3 has no out-edge), 0->4, 4->0, 5->3
    adj = {0: [1, 4], 1: [2], 2: [3], 3: [], 4: [0], 5: [3]}
    rowptr = [0]
    col = []
    for v in range(6):
        col += adj[v]
        rowptr.append(len(col))
    return np.array(rowptr, np.uint64), np.array(col, np.uint32)


def test_ppr_dangling_absorbs():
    rowptr, col = _dangling_graph()
    w = 100000
    p = oracle.L.ltor_threshold(float(np.float32(0.15))) / 2 ** 32
    counts, _, _ = oracle.ppr(rowptr, col, w, source=0)
    assert counts.sum() == w
    r = _ppr_expectation(rowptr, col, 0, p) * w
    assert np.all(np.abs(counts - r) <= 6 * np.sqrt(r) + 1)
    assert counts[5] == 0


def test_genericwalk_distribution_and_stops():
    rowptr, col = _dangling_graph()
    V = 6
    w, l = 120000, 4
    ends, moves, st = oracle.genericwalk(rowptr, col, w, l)
    deg = rowptr[1:] - rowptr[:-1]
    assert np.all(moves <= l)
    assert np.all(deg[ends[moves < l]] == 0)  # early stop only at a sink
    assert st[0] == moves.sum()
    P = _dense_P(rowptr, col)
    P[3, 3] = 1.0
    x = np.full(V, 1.0 / V) @ np.linalg.matrix_power(P, l)
    cnt = np.bincount(ends, minlength=V)
    assert np.all(np.abs(cnt - w * x) <= 6 * np.sqrt(w * x) + 1)


def test_walks_deterministic_and_thread_independent():
    rowptr, col = _small_graph()
    w = 5000
    a = oracle.pagerank(rowptr, col, w, 8, threads=1, endpoints=True)
    b = oracle.pagerank(rowptr, col, w, 8, threads=4, endpoints=True)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    # an id range is a slice of the full run
    c = oracle.pagerank(rowptr, col, w, 8, id_begin=1000, id_end=3000, counts=False, endpoints=True)
    assert np.array_equal(c[1], a[1][1000:3000])


def test_runs_extend_walker_ids():
    rowptr, col = _small_graph()
    w = 3000
    two = oracle.genericwalk(rowptr, col, w, 5, runs=2)[0]
    second = oracle.genericwalk(rowptr, col, w, 5, id_begin=w, id_end=2 * w)[0]
    assert np.array_equal(two[w:], second)


def test_ppr_batch_equals_single_source_slices():
    rowptr, col = _small_graph()
    sources = [0, 5, 9]
    qc = [1000, 20, 333]
    ends, _ = oracle.ppr_batch(rowptr, col, sources, qc)
    off = 0
    for s, n in zip(sources, qc):
        _, e, _ = oracle.ppr(rowptr, col, 0, source=s, id_begin=off, id_end=off + n, counts=False, endpoints=True)
        assert np.array_equal(ends[off:off + n], e)
        off += n


def test_partition_greedy_invariants():
    rng = np.random.default_rng(3)
    for trial in range(20):
        V = int(rng.integers(1, 400))
        deg = rng.integers(0, 50, V)
        rowptr = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint64)
        S = int(rng.integers(8 * 2 + 4 * 50, 4000))
        no, eo = oracle.partition(rowptr, S)
        assert no[0] == 0 and no[-1] == V and np.all(np.diff(no.astype(np.int64)) > 0)
        assert np.array_equal(eo, rowptr[no])
        for k in range(len(no) - 1):
            a, b = int(no[k]), int(no[k + 1])
            size = (b - a + 1) * 8 + int(rowptr[b] - rowptr[a]) * 4
            assert size <= S
            if b < V:  # maximal: one more vertex would overflow
                assert (b + 1 - a + 1) * 8 + int(rowptr[b + 1] - rowptr[a]) * 4 > S


def test_partition_oversize_vertex():
    rowptr = np.array([0, 3, 1003, 1005], np.uint64)
    with pytest.raises(ValueError):
        oracle.partition(rowptr, 1000)


def test_find_partition_bruteforce():
    no = np.array([0, 3, 3, 10, 11, 40], np.uint32)
    for v in range(40):
        expect = max(k for k in range(len(no) - 1) if no[k] <= v)
        assert oracle.find_partition(no, v) == expect


def test_genericwalk_paths_follow_edges_and_match_walk():
    """Paths restate the same walk as ltor_genericwalk: column 0 is the Philox
    start, every next column a neighbour of the previous one, the last valid
    column the endpoint, the valid columns one more than the moves; columns
    after an early stop are 0xFFFFFFFF."""
    rng = np.random.default_rng(4)
    V = 500
    deg = rng.integers(0, 5, V)
    deg[rng.random(V) < 0.2] = 0
    rowptr = np.concatenate([[0], np.cumsum(deg)]).astype(np.uint64)
    col = rng.integers(0, V, int(rowptr[-1])).astype(np.uint32)
    w, l = 3000, 6
    paths = oracle.genericwalk_paths(rowptr, col, w, l, runs=2)
    ends, moves, _ = oracle.genericwalk(rowptr, col, w, l, runs=2)
    assert paths.shape == (2 * w, l + 1)
    for i in range(2 * w):
        p = paths[i]
        valid = int(moves[i]) + 1
        assert np.all(p[valid:] == 0xFFFFFFFF) and np.all(p[:valid] < V)
        assert p[valid - 1] == ends[i]
        x = oracle.philox([i & 0xFFFFFFFF, i >> 32, 0, 0x53545254], [0x4C54524B, 0])
        assert p[0] == (((x[0] << 32) | x[1]) * V) >> 64
        for k in range(valid - 1):
            a = int(p[k])
            assert p[k + 1] in col[rowptr[a]:rowptr[a + 1]]
        if valid <= l:  # stopped early: at a vertex without neighbours
            assert rowptr[p[valid - 1] + 1] == rowptr[p[valid - 1]]
