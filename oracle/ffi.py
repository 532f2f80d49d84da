"""ctypes bindings of the CPU oracle (oracle/lt_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the engine (lighttraffic_amd).
See oracle/lt_oracle.c for which reference file:line each function restates.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build", "liblt_oracle.so")

_vp = ctypes.c_void_p
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load():
    if not os.path.exists(LIB):
        build()
    L = ctypes.CDLL(LIB)

    def sig(name, res, *args):
        f = getattr(L, name)
        f.restype = res
        f.argtypes = list(args)

    sig("ltor_philox", None, _vp, _vp, _vp)
    sig("ltor_threshold", _u32, ctypes.c_double)
    sig("ltor_teleport_threshold", _u32)
    sig("ltor_pagerank", ctypes.c_int, _vp, _vp, _u32, _u32, _u64, _u64, _u64, _vp, _vp, ctypes.c_int, _vp)
    sig("ltor_ppr", ctypes.c_int, _vp, _vp, _u32, _u32, ctypes.c_float, _u64, _u64, _u64, _vp, _vp, ctypes.c_int,
        _vp)
    sig("ltor_ppr_batch", ctypes.c_int, _vp, _vp, _u32, _u32, _vp, _vp, ctypes.c_float, _u64, _vp, ctypes.c_int,
        _vp)
    sig("ltor_genericwalk", ctypes.c_int, _vp, _vp, _u32, _u32, _u64, _u64, _u64, _vp, _vp, ctypes.c_int, _vp)
    sig("ltor_genericwalk_paths", ctypes.c_int, _vp, _vp, _u32, _u32, _u64, _u64, _u64, _vp, ctypes.c_int)
    sig("ltor_rmat_edges", ctypes.c_int, _u32, _u64, ctypes.c_double, ctypes.c_double, ctypes.c_double, _u64, _vp,
        _vp, ctypes.c_int)
    sig("ltor_convert", ctypes.c_int, _u64, _vp, _vp, ctypes.POINTER(_u32), ctypes.POINTER(_u64),
        ctypes.POINTER(_vp), ctypes.POINTER(_vp))
    sig("ltor_free", None, _vp)
    sig("ltor_partition", ctypes.c_int64, _vp, _u32, _u64, _vp, _vp, _u64, ctypes.POINTER(_u32))
    sig("ltor_find_partition", _u32, _vp, _u32, _u32)
    sig("ltor_greedy_next", ctypes.c_int32, _u32, _vp, ctypes.c_int32, ctypes.c_int32)
    sig("ltor_evict_min", ctypes.c_int32, _u32, _vp, _vp, ctypes.c_int32)
    sig("ltor_computed_partition", ctypes.c_int32, _u32, _vp, _vp, ctypes.c_int32, _u64)
    return L


L = _load()


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    L.ltor_philox(c, k, o)
    return list(o)


def _arrs(rowptr, col):
    return (np.ascontiguousarray(rowptr, dtype=np.uint64), np.ascontiguousarray(col, dtype=np.uint32))


def pagerank(rowptr, col, w, l, runs=1, seed=0x4C54524B, id_begin=0, id_end=None, threads=0, counts=True,
             endpoints=False):
    rowptr, col = _arrs(rowptr, col)
    V = len(rowptr) - 1
    if id_end is None:
        id_end = w * runs
    c = np.zeros(V, np.uint32) if counts else None
    e = np.zeros(id_end - id_begin, np.uint32) if endpoints else None
    st = np.zeros(3, np.uint64)
    rc = L.ltor_pagerank(rowptr.ctypes.data, col.ctypes.data, V, l, id_begin, id_end, seed,
                         None if c is None else c.ctypes.data, None if e is None else e.ctypes.data, threads,
                         st.ctypes.data)
    assert rc == 0
    return c, e, st


def ppr(rowptr, col, w, source=0, prob=0.15, seed=0x4C54524B, id_begin=0, id_end=None, threads=0, counts=True,
        endpoints=False):
    rowptr, col = _arrs(rowptr, col)
    V = len(rowptr) - 1
    if id_end is None:
        id_end = w
    c = np.zeros(V, np.uint32) if counts else None
    e = np.zeros(id_end - id_begin, np.uint32) if endpoints else None
    st = np.zeros(3, np.uint64)
    rc = L.ltor_ppr(rowptr.ctypes.data, col.ctypes.data, V, source, prob, id_begin, id_end, seed,
                    None if c is None else c.ctypes.data, None if e is None else e.ctypes.data, threads,
                    st.ctypes.data)
    assert rc == 0
    return c, e, st


def ppr_batch(rowptr, col, sources, qcounts, prob=0.15, seed=0x4C54524B, threads=0):
    rowptr, col = _arrs(rowptr, col)
    V = len(rowptr) - 1
    sources = np.ascontiguousarray(sources, dtype=np.uint32)
    qcounts = np.ascontiguousarray(qcounts, dtype=np.uint64)
    e = np.zeros(int(qcounts.sum()), np.uint32)
    st = np.zeros(3, np.uint64)
    rc = L.ltor_ppr_batch(rowptr.ctypes.data, col.ctypes.data, V, len(sources), sources.ctypes.data,
                          qcounts.ctypes.data, prob, seed, e.ctypes.data, threads, st.ctypes.data)
    assert rc == 0
    return e, st


def genericwalk(rowptr, col, w, l, runs=1, seed=0x4C54524B, id_begin=0, id_end=None, threads=0):
    rowptr, col = _arrs(rowptr, col)
    V = len(rowptr) - 1
    if id_end is None:
        id_end = w * runs
    e = np.zeros(id_end - id_begin, np.uint32)
    m = np.zeros(id_end - id_begin, np.uint32)
    st = np.zeros(3, np.uint64)
    rc = L.ltor_genericwalk(rowptr.ctypes.data, col.ctypes.data, V, l, id_begin, id_end, seed, e.ctypes.data,
                            m.ctypes.data, threads, st.ctypes.data)
    assert rc == 0
    return e, m, st


def genericwalk_paths(rowptr, col, w, l, runs=1, seed=0x4C54524B, threads=0):
    """u32[w * runs, l + 1]: vertex after k moves at column k, 0xFFFFFFFF after an early stop."""
    rowptr, col = _arrs(rowptr, col)
    V = len(rowptr) - 1
    p = np.zeros((w * runs, l + 1), np.uint32)
    rc = L.ltor_genericwalk_paths(rowptr.ctypes.data, col.ctypes.data, V, l, 0, w * runs, seed, p.ctypes.data, threads)
    assert rc == 0
    return p


def rmat_edges(scale, edge_factor, a=0.57, b=0.19, c=0.19, seed=None, threads=0):
    if seed is None:
        seed = 0x5EED0000 + scale
    n = edge_factor << scale
    src = np.zeros(n, np.uint32)
    dst = np.zeros(n, np.uint32)
    rc = L.ltor_rmat_edges(scale, n, a, b, c, seed, src.ctypes.data, dst.ctypes.data, threads)
    assert rc == 0
    return src, dst


def convert(src, dst):
    """Edge list -> (rowptr, col) by the rules of tools/convert_from_txt.cc."""
    s = np.ascontiguousarray(src, dtype=np.uint64)
    d = np.ascontiguousarray(dst, dtype=np.uint64)
    V = _u32()
    E = _u64()
    rp = _vp()
    cp = _vp()
    rc = L.ltor_convert(len(s), s.ctypes.data, d.ctypes.data, ctypes.byref(V), ctypes.byref(E), ctypes.byref(rp),
                        ctypes.byref(cp))
    assert rc == 0
    rowptr = np.ctypeslib.as_array(ctypes.cast(rp, ctypes.POINTER(_u64)), (V.value + 1,)).copy()
    col = np.ctypeslib.as_array(ctypes.cast(cp, ctypes.POINTER(_u32)), (max(E.value, 1),))[:E.value].copy()
    L.ltor_free(rp)
    L.ltor_free(cp)
    return rowptr, col


def rmat_graph(scale, edge_factor, seed=None, threads=0):
    src, dst = rmat_edges(scale, edge_factor, seed=seed, threads=threads)
    return convert(src, dst)


def bcsr_bytes(rowptr, col):
    V = len(rowptr) - 1
    return (np.uint32(V).tobytes() + np.uint64(len(col)).tobytes() + np.asarray(rowptr, np.uint64).tobytes() +
            np.asarray(col, np.uint32).tobytes())


def read_bcsr(path):
    with open(path, "rb") as f:
        raw = f.read()
    V = int(np.frombuffer(raw[:4], np.uint32)[0])
    E = int(np.frombuffer(raw[4:12], np.uint64)[0])
    rowptr = np.frombuffer(raw[12:12 + 8 * (V + 1)], np.uint64).copy()
    col = np.frombuffer(raw[12 + 8 * (V + 1):], np.uint32).copy()
    assert len(col) == E
    return rowptr, col


def partition(rowptr, max_bytes):
    rowptr = np.ascontiguousarray(rowptr, dtype=np.uint64)
    V = len(rowptr) - 1
    cap = V + 2
    no = np.zeros(cap, np.uint32)
    eo = np.zeros(cap, np.uint64)
    bad = _u32()
    P = L.ltor_partition(rowptr.ctypes.data, V, max_bytes, no.ctypes.data, eo.ctypes.data, cap, ctypes.byref(bad))
    if P == -1:
        raise ValueError(f"vertex {bad.value} does not fit")
    assert P > 0
    return no[:P + 1].copy(), eo[:P + 1].copy()


def find_partition(node_off, v):
    no = np.ascontiguousarray(node_off, dtype=np.uint32)
    return L.ltor_find_partition(no.ctypes.data, len(no) - 1, v)


def greedy_next(counts, prev, sync):
    c = np.ascontiguousarray(counts, dtype=np.uint64)
    return L.ltor_greedy_next(len(c), c.ctypes.data, prev, int(sync))


def evict_min(counts, resident, except_):
    c = np.ascontiguousarray(counts, dtype=np.uint64)
    r = np.ascontiguousarray(resident, dtype=np.uint8)
    return L.ltor_evict_min(len(c), c.ctypes.data, r.ctypes.data, except_)


def computed_partition(counts, resident, except_, page):
    c = np.ascontiguousarray(counts, dtype=np.uint64)
    r = np.ascontiguousarray(resident, dtype=np.uint8)
    return L.ltor_computed_partition(len(c), c.ctypes.data, r.ctypes.data, except_, page)
