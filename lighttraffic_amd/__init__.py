"""Python bindings of liblighttraffic.so, the MI355X random-walk engine.

The engine itself is C++/HIP behind the C ABI declared in
include/lighttraffic.h; this module only loads that library with ctypes and
wraps its handles so tests and the benchmark can drive it.  It mirrors the
operator surface of the reference programs (app/pagerank.cu, app/ppr.cu,
app/genericwalk.cu and the flags of src/Config.h:38-89): a graph in bcsr
form, a materialised partitioning, and one run call per walk kind.

Importing never falls back to another implementation: if the shared library
has not been built, the import raises.
"""
import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "liblighttraffic.so")
BIN_DIR = os.path.join(_HERE, "bin")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                      "or `make -C lighttraffic_amd/csrc`")

_lib = ctypes.CDLL(LIB_PATH)

LT_OK = 0
STATUS = {0: "LT_OK", 1: "LT_E_IO", 2: "LT_E_FORMAT", 3: "LT_E_OVERSIZE_VERTEX", 4: "LT_E_NOMEM", 5: "LT_E_HIP",
          6: "LT_E_ARG", 7: "LT_E_NODEVICE", 8: "LT_E_INTERNAL"}
SCHED = {"greedy": 0, "fifo": 1, "learned": 2}
SPAN = {"partition": 0, "resident": 1}

_vp = ctypes.c_void_p
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_i32 = ctypes.c_int32


class LTError(RuntimeError):
    def __init__(self, status, where):
        self.status = status
        detail = _lib.lt_last_error().decode(errors="replace")
        super().__init__(f"{where}: {STATUS.get(status, status)} ({detail})")


class Config(ctypes.Structure):
    """lt_config (include/lighttraffic.h)."""
    _fields_ = [("num_device_partitions", _u32), ("num_device_walkers", _u64), ("zerocopy_threshold", _u64),
                ("seed", _u64), ("policy", _i32), ("pipeline", _i32), ("span", _i32), ("event_timers", _i32),
                ("record_trace", _i32), ("prefetch_depth", _i32), ("prefetch_stick_pct", _i32),
                ("device_schedule", _i32), ("slice_chunks", _i32), ("zc_batch", _i32),
                ("rowptr_resident", _i32), ("pack_partitions", _i32), ("prefetch_plain", _i32),
                ("evict_shortlist", _i32), ("walkers_per_thread", _i32), ("resident_batch", _i32),
                ("learned_next", _i32), ("append_mode", _i32)]


class Stats(ctypes.Structure):
    """lt_stats (include/lighttraffic.h)."""
    _fields_ = [("running_ms", ctypes.c_double), ("iterations", _u64), ("explicit_iterations", _u64),
                ("graph_loading_ms", ctypes.c_double), ("graph_loading_calls", _u64),
                ("computing_ms", ctypes.c_double), ("computing_calls", _u64),
                ("zerocopy_ms", ctypes.c_double), ("zerocopy_calls", _u64),
                ("walker_loading_ms", ctypes.c_double), ("walker_loading_calls", _u64),
                ("walker_evicting_ms", ctypes.c_double), ("walker_evicting_calls", _u64),
                ("h2d_bytes_explicit", _u64), ("h2d_bytes_zerocopy", _u64), ("moves", _u64),
                ("teleports", _u64), ("visits", _u64), ("residencies", _u64), ("zerocopy_moves", _u64),
                ("nominal_steps", _u64), ("kernel_launches", _u64), ("walk_kernel_ms", ctypes.c_double),
                ("learned_decisions", _u64), ("fallback_decisions", _u64), ("prefetch_issued", _u64),
                ("prefetch_hits", _u64), ("h2d_bytes_unpacked", _u64), ("packed_image_bytes", _u64),
                ("packed_col_bytes", _u64), ("cache_slots", _u64), ("zerocopy_launches", _u64),
                ("walk_kernel_zc_ms", ctypes.c_double), ("resident_batched", _u64), ("next_overrides", _u64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


class TraceRec(ctypes.Structure):
    _fields_ = [("partition", _i32), ("kind", _i32), ("loaded", _i32), ("evicted", _i32), ("walkers", _u64)]


def _sig(name, restype, *argtypes):
    f = getattr(_lib, name)
    f.restype = restype
    f.argtypes = list(argtypes)
    return f


_sig("lt_status_str", ctypes.c_char_p, ctypes.c_int)
_sig("lt_last_error", ctypes.c_char_p)
_sig("lt_config_default", None, ctypes.POINTER(Config))
_sig("lt_graph_open", ctypes.c_int, ctypes.c_char_p, ctypes.POINTER(_vp))
_sig("lt_graph_from_csr", ctypes.c_int, _u32, _u64, _vp, _vp, ctypes.POINTER(_vp))
_sig("lt_graph_generate_rmat", ctypes.c_int, _u32, _u32, ctypes.c_double, ctypes.c_double, ctypes.c_double, _u64,
     ctypes.POINTER(_vp))
_sig("lt_graph_save", ctypes.c_int, _vp, ctypes.c_char_p)
_sig("lt_graph_num_node", _u32, _vp)
_sig("lt_graph_num_edge", _u64, _vp)
_sig("lt_graph_rowptr", _vp, _vp)
_sig("lt_graph_col", _vp, _vp)
_sig("lt_graph_free", None, _vp)
_sig("lt_partition_build", ctypes.c_int, _vp, _u64, ctypes.POINTER(_vp))
_sig("lt_partition_save", ctypes.c_int, _vp, ctypes.c_char_p)
_sig("lt_partition_load", ctypes.c_int, ctypes.c_char_p, ctypes.POINTER(_vp))
_sig("lt_partition_from_offsets", ctypes.c_int, _u32, _u64, _vp, _vp, ctypes.POINTER(_vp))
_sig("lt_partition_count", _u32, _vp)
_sig("lt_partition_bytes", _u64, _vp)
_sig("lt_partition_node_offsets", _vp, _vp)
_sig("lt_partition_edge_offsets", _vp, _vp)
_sig("lt_partition_find", _u32, _vp, _u32)
_sig("lt_partition_free", None, _vp)
_sig("lt_partition_pack", ctypes.c_int, _vp, _vp, _u32, _vp, _u64, ctypes.POINTER(_u64))
_sig("lt_partition_pack_blocks", ctypes.c_int, _vp, _vp, _u32, _vp, _u64, ctypes.POINTER(_u64))
_sig("lt_partition_build_blocks", ctypes.c_int, _vp, _u64, ctypes.POINTER(_vp))
_sig("lt_partition_slot_layout", ctypes.c_int, _vp, _vp, ctypes.c_int32, _u32, ctypes.POINTER(_u64),
     ctypes.POINTER(_u32))
_sig("lt_graph_drop_packed", None, _vp)
_sig("lt_engine_fetch_partition", ctypes.c_int, _vp, _u32, _vp, _vp)
_sig("lt_engine_create", ctypes.c_int, _vp, _vp, ctypes.POINTER(Config), ctypes.POINTER(_vp))
_sig("lt_engine_destroy", None, _vp)
_sig("lt_run_pagerank", ctypes.c_int, _vp, _u64, _u32, _u32, _vp, _vp)
_sig("lt_run_ppr", ctypes.c_int, _vp, _u64, _u32, ctypes.c_float, _vp, _vp)
_sig("lt_run_genericwalk", ctypes.c_int, _vp, _u64, _u32, _u32, _vp, _vp)
_sig("lt_run_genericwalk_paths", ctypes.c_int, _vp, _u64, _u32, _u32, _vp)
_sig("lt_run_ppr_batch", ctypes.c_int, _vp, _u32, _vp, _vp, ctypes.c_float, _vp)
_sig("lt_engine_stats", ctypes.c_int, _vp, ctypes.POINTER(Stats))
_sig("lt_make_ppr_queries", ctypes.c_int, _vp, _u32, _u64, _u64, _u64, _vp, _vp)
_sig("lt_engine_trace_len", _u64, _vp)
_sig("lt_engine_trace", ctypes.c_int, _vp, ctypes.POINTER(TraceRec), _u64)
_sig("lt_engine_trace_state", ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _u64)
_sig("lt_engine_set_scorer", ctypes.c_int, _vp, _vp)
_sig("lt_scorer_create", ctypes.c_int, _vp, _u64, ctypes.POINTER(_vp))
_sig("lt_scorer_free", None, _vp)
_sig("lt_scorer_num_params", _u64)
_sig("lt_scorer_add_trace", ctypes.c_int, _vp, _vp, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(_u64))
_sig("lt_scorer_fit", ctypes.c_int, _vp, _u32, _u32, ctypes.c_float, _u64, _vp)
_sig("lt_scorer_train", ctypes.c_int, _vp, _u64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _u32, _u32,
     ctypes.c_float, _u64, _vp)
_sig("lt_scorer_loss_grad", ctypes.c_int, _vp, _u64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp)
_sig("lt_scorer_infer", ctypes.c_int, _vp, _vp, _vp, _vp, _vp, ctypes.c_int32, ctypes.c_float, _vp)
_sig("lt_scorer_infer_noalloc", ctypes.c_int, _vp, _vp, _vp, _vp, _vp, ctypes.c_int32, ctypes.c_float, _vp)
_sig("lt_scorer_get_params", ctypes.c_int, _vp, _vp)
_sig("lt_scorer_set_params", ctypes.c_int, _vp, _vp)
_sig("lt_scorer_save", ctypes.c_int, _vp, ctypes.c_char_p)
_sig("lt_scorer_load", ctypes.c_int, _vp, ctypes.c_char_p)
_sig("lt_sched_greedy_next", _i32, _u32, _vp, _i32, _i32)
_sig("lt_sched_evict_min", _i32, _u32, _vp, _vp, _i32)
_sig("lt_sched_computed_partition", _i32, _u32, _vp, _vp, _i32, _u64)
_sig("lt_sched_belady", ctypes.c_int, _vp, _u64, _u32, _u32, _vp, _vp, ctypes.POINTER(_u64), ctypes.POINTER(_u64),
     _vp)
_sig("lt_device_info", ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_int),
     ctypes.POINTER(_u64))
_sig("lt_device_copy_bandwidth", ctypes.c_int, _u64, _u32, ctypes.POINTER(ctypes.c_double))
_sig("lt_device_h2d_bandwidth", ctypes.c_int, _u64, _u32, ctypes.POINTER(ctypes.c_double))


def lib():
    return _lib


def _check(status, where):
    if status != LT_OK:
        raise LTError(status, where)


def _ptr(a):
    return None if a is None else a.ctypes.data


def default_config(**kw):
    c = Config()
    _lib.lt_config_default(ctypes.byref(c))
    names = {f for f, _ in Config._fields_ if f != "reserved"}
    for k, v in kw.items():
        if k not in names:
            raise TypeError(f"unknown engine option {k!r}")
        if k == "policy" and isinstance(v, str):
            v = SCHED[v]
        if k == "span" and isinstance(v, str):
            v = SPAN[v]
        setattr(c, k, v)
    return c


def device_info():
    name = ctypes.create_string_buffer(256)
    ncu = ctypes.c_int()
    hbm = _u64()
    _check(_lib.lt_device_info(name, 256, ctypes.byref(ncu), ctypes.byref(hbm)), "lt_device_info")
    return {"name": name.value.decode(), "num_cu": ncu.value, "hbm_bytes": hbm.value}


def device_copy_bandwidth(nbytes, reps=5):
    """HBM copy bandwidth (read + written bytes per second) of a plain copy kernel."""
    gbps = ctypes.c_double()
    _check(_lib.lt_device_copy_bandwidth(nbytes, reps, ctypes.byref(gbps)), "lt_device_copy_bandwidth")
    return gbps.value


def device_h2d_bandwidth(nbytes, reps=5):
    """Pinned host -> HBM DMA rate in GB/s (hipMemcpyAsync of nbytes, reps times)."""
    gbps = ctypes.c_double()
    _check(_lib.lt_device_h2d_bandwidth(nbytes, reps, ctypes.byref(gbps)), "lt_device_h2d_bandwidth")
    return gbps.value


class Graph:
    """A bcsr graph (u32 vertex ids, u64 edge offsets) owned by the library."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def open(cls, path):
        h = _vp()
        _check(_lib.lt_graph_open(path.encode(), ctypes.byref(h)), f"lt_graph_open({path})")
        return cls(h)

    @classmethod
    def from_csr(cls, rowptr, col):
        rowptr = np.ascontiguousarray(rowptr, dtype=np.uint64)
        col = np.ascontiguousarray(col, dtype=np.uint32)
        h = _vp()
        _check(_lib.lt_graph_from_csr(len(rowptr) - 1, len(col), rowptr.ctypes.data, col.ctypes.data,
                                      ctypes.byref(h)), "lt_graph_from_csr")
        return cls(h)

    @classmethod
    def rmat(cls, scale, edge_factor, a=0.57, b=0.19, c=0.19, seed=None):
        if seed is None:
            seed = 0x5EED0000 + scale
        h = _vp()
        _check(_lib.lt_graph_generate_rmat(scale, edge_factor, a, b, c, seed, ctypes.byref(h)),
               "lt_graph_generate_rmat")
        return cls(h)

    def save(self, path):
        _check(_lib.lt_graph_save(self._h, path.encode()), f"lt_graph_save({path})")

    def drop_packed(self):
        """Free the packed partition images cached on this graph (engines using
        one keep it until they are closed)."""
        _lib.lt_graph_drop_packed(self._h)

    def ppr_queries(self, nq, lo=1000, hi=1000000, seed=0x51525953):
        """Batched-PPR query table: sources proportional to degree, log-uniform counts."""
        src = np.zeros(nq, np.uint32)
        cnt = np.zeros(nq, np.uint64)
        _check(_lib.lt_make_ppr_queries(self._h, nq, lo, hi, seed, src.ctypes.data, cnt.ctypes.data),
               "lt_make_ppr_queries")
        return src, cnt

    @property
    def num_node(self):
        return _lib.lt_graph_num_node(self._h)

    @property
    def num_edge(self):
        return _lib.lt_graph_num_edge(self._h)

    def rowptr(self):
        """Zero-copy numpy view of rowptr (valid while the graph lives)."""
        n = self.num_node + 1
        return np.ctypeslib.as_array(ctypes.cast(_lib.lt_graph_rowptr(self._h), ctypes.POINTER(_u64)), (n,))

    def col(self):
        n = max(self.num_edge, 1)
        return np.ctypeslib.as_array(ctypes.cast(_lib.lt_graph_col(self._h), ctypes.POINTER(_u32)), (n,))[
            :self.num_edge]

    def close(self):
        if self._h and _lib is not None:
            _lib.lt_graph_free(self._h)
        self._h = None

    def __del__(self):
        self.close()


class Partition:
    """Contiguous vertex-range partitions (the .part file)."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def build(cls, graph, max_bytes):
        h = _vp()
        _check(_lib.lt_partition_build(graph._h, max_bytes, ctypes.byref(h)), "lt_partition_build")
        return cls(h)

    @classmethod
    def build_blocks(cls, graph, slot_bytes):
        """Partitions whose rowptr slice + block image fit slot_bytes each."""
        h = _vp()
        _check(_lib.lt_partition_build_blocks(graph._h, slot_bytes, ctypes.byref(h)), "lt_partition_build_blocks")
        return cls(h)

    @classmethod
    def load(cls, graph_path):
        h = _vp()
        _check(_lib.lt_partition_load(graph_path.encode(), ctypes.byref(h)), "lt_partition_load")
        return cls(h)

    def save(self, graph_path):
        _check(_lib.lt_partition_save(self._h, graph_path.encode()), "lt_partition_save")

    @property
    def count(self):
        return _lib.lt_partition_count(self._h)

    @property
    def bytes(self):
        return _lib.lt_partition_bytes(self._h)

    def node_offsets(self):
        n = self.count + 1
        return np.ctypeslib.as_array(ctypes.cast(_lib.lt_partition_node_offsets(self._h), ctypes.POINTER(_u32)),
                                     (n,)).copy()

    def edge_offsets(self):
        n = self.count + 1
        return np.ctypeslib.as_array(ctypes.cast(_lib.lt_partition_edge_offsets(self._h), ctypes.POINTER(_u64)),
                                     (n,)).copy()

    def find(self, v):
        return _lib.lt_partition_find(self._h, v)

    def pack(self, graph, part, blocks=False):
        """Packed image (u8 array) of partition `part`'s column slice, the
        form an engine with pack_partitions=1 copies to HBM (blocks=True: the
        block image that pack_partitions=2 keeps in HBM)."""
        name = "lt_partition_pack_blocks" if blocks else "lt_partition_pack"
        fn = getattr(_lib, name)
        n = _u64()
        _check(fn(graph._h, self._h, part, None, 0, ctypes.byref(n)), name)
        out = np.zeros(n.value, np.uint8)
        _check(fn(graph._h, self._h, part, out.ctypes.data, n.value, ctypes.byref(n)), name)
        return out

    def slot_layout(self, graph, pack_partitions=0, num_device_partitions=0):
        """(slot bytes, slots) of the HBM cache an engine with these settings
        gets (before rowptr_resident takes its share)."""
        sb, n = _u64(), _u32()
        _check(_lib.lt_partition_slot_layout(graph._h, self._h, pack_partitions, num_device_partitions,
                                             ctypes.byref(sb), ctypes.byref(n)), "lt_partition_slot_layout")
        return sb.value, n.value

    def close(self):
        if self._h and _lib is not None:
            _lib.lt_partition_free(self._h)
        self._h = None

    def __del__(self):
        self.close()


class Engine:
    """GPU engine bound to one graph and partitioning (lt_engine)."""

    def __init__(self, graph, partition, config=None, **kw):
        cfg = config if config is not None else default_config(**kw)
        h = _vp()
        _check(_lib.lt_engine_create(graph._h, partition._h, ctypes.byref(cfg), ctypes.byref(h)),
               "lt_engine_create")
        self._h = h
        self.graph = graph
        self.partition = partition
        self.config = cfg

    def pagerank(self, w, l=10, runs=1, endpoints=False):
        counts = np.zeros(self.graph.num_node, np.uint32)
        ep = np.zeros(w * runs, np.uint32) if endpoints else None
        _check(_lib.lt_run_pagerank(self._h, w, l, runs, counts.ctypes.data, _ptr(ep)), "lt_run_pagerank")
        return (counts, ep) if endpoints else counts

    def ppr(self, w, source=0, prob=0.15, endpoints=False):
        counts = np.zeros(self.graph.num_node, np.uint32)
        ep = np.zeros(w, np.uint32) if endpoints else None
        _check(_lib.lt_run_ppr(self._h, w, source, prob, counts.ctypes.data, _ptr(ep)), "lt_run_ppr")
        return (counts, ep) if endpoints else counts

    def genericwalk(self, w, l=10, runs=1, endpoints=True, moves=False):
        ep = np.zeros(w * runs, np.uint32) if endpoints else None
        mv = np.zeros(w * runs, np.uint32) if moves else None
        _check(_lib.lt_run_genericwalk(self._h, w, l, runs, _ptr(ep), _ptr(mv)), "lt_run_genericwalk")
        return ep, mv

    def genericwalk_paths(self, w, l=10, runs=1):
        """u32[w * runs, l + 1]: column k is the vertex after k moves, 0xFFFFFFFF after an early stop."""
        paths = np.zeros((w * runs, l + 1), np.uint32)
        _check(_lib.lt_run_genericwalk_paths(self._h, w, l, runs, paths.ctypes.data), "lt_run_genericwalk_paths")
        return paths

    def ppr_batch(self, sources, counts, prob=0.15, endpoints=True):
        sources = np.ascontiguousarray(sources, dtype=np.uint32)
        counts = np.ascontiguousarray(counts, dtype=np.uint64)
        ep = np.zeros(int(counts.sum()), np.uint32) if endpoints else None
        _check(_lib.lt_run_ppr_batch(self._h, len(sources), sources.ctypes.data, counts.ctypes.data, prob,
                                     _ptr(ep)), "lt_run_ppr_batch")
        return ep

    def fetch_partition(self, part, node_offsets, edge_offsets):
        """(rowptr u64[n+1], col u32[m]) of partition `part` as the engine's
        partition copy writes them into an HBM slot."""
        n = int(node_offsets[part + 1]) - int(node_offsets[part])
        m = int(edge_offsets[part + 1]) - int(edge_offsets[part])
        rowptr = np.zeros(n + 1, np.uint64)
        col = np.zeros(max(m, 1), np.uint32)
        _check(_lib.lt_engine_fetch_partition(self._h, part, rowptr.ctypes.data, col.ctypes.data),
               "lt_engine_fetch_partition")
        return rowptr, col[:m]

    def stats(self):
        s = Stats()
        _check(_lib.lt_engine_stats(self._h, ctypes.byref(s)), "lt_engine_stats")
        return s.as_dict()

    def trace(self):
        n = _lib.lt_engine_trace_len(self._h)
        buf = (TraceRec * max(n, 1))()
        _check(_lib.lt_engine_trace(self._h, buf, n), "lt_engine_trace")
        return [(r.partition, r.kind, r.loaded, r.evicted, r.walkers) for r in buf[:n]]

    def trace_state(self):
        """(counts (len,P) u32, resident (len,P) u8, age (len,P) u32, recent (len,8) i32) seen by each
        traced decision."""
        n = _lib.lt_engine_trace_len(self._h)
        P = self.partition.count
        c = np.zeros((n, P), np.uint32)
        r = np.zeros((n, P), np.uint8)
        a = np.zeros((n, P), np.uint32)
        rc = np.zeros((n, 8), np.int32)
        _check(_lib.lt_engine_trace_state(self._h, c.ctypes.data, r.ctypes.data, a.ctypes.data, rc.ctypes.data, n),
               "lt_engine_trace_state")
        return c, r, a, rc

    def set_scorer(self, scorer):
        _check(_lib.lt_engine_set_scorer(self._h, None if scorer is None else scorer._h), "lt_engine_set_scorer")
        self._scorer = scorer

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.lt_engine_destroy(self._h)
        self._h = None

    def __del__(self):
        self.close()


def sched_greedy_next(counts, prev, sync):
    c = np.ascontiguousarray(counts, dtype=np.uint64)
    return _lib.lt_sched_greedy_next(len(c), c.ctypes.data, prev, int(sync))


def sched_evict_min(counts, resident, except_):
    c = np.ascontiguousarray(counts, dtype=np.uint64)
    r = np.ascontiguousarray(resident, dtype=np.uint8)
    return _lib.lt_sched_evict_min(len(c), c.ctypes.data, r.ctypes.data, except_)


def sched_computed_partition(counts, resident, except_, page):
    c = np.ascontiguousarray(counts, dtype=np.uint64)
    r = np.ascontiguousarray(resident, dtype=np.uint8)
    return _lib.lt_sched_computed_partition(len(c), c.ctypes.data, r.ctypes.data, except_, page)


def sched_belady(seq, P, slots, init_resident=None, part_bytes=None):
    """Belady's MIN on demand sequence `seq`: (loads, bytes, victims[i] or -1)."""
    q = np.ascontiguousarray(seq, dtype=np.int32)
    r = None if init_resident is None else np.ascontiguousarray(init_resident, dtype=np.uint8)
    b = None if part_bytes is None else np.ascontiguousarray(part_bytes, dtype=np.uint64)
    victims = np.zeros(max(len(q), 1), np.int32)
    loads, nbytes = _u64(), _u64()
    _check(_lib.lt_sched_belady(q.ctypes.data, len(q), P, slots, _ptr(r), _ptr(b), ctypes.byref(loads),
                                ctypes.byref(nbytes), victims.ctypes.data), "lt_sched_belady")
    return loads.value, nbytes.value, victims[:len(q)]


class Scorer:
    """Learned partition scorer: per-partition BF16 MLP trained by HIP kernels."""

    NUM_PARAMS = int(_lib.lt_scorer_num_params())

    def __init__(self, partition, seed=1):
        h = _vp()
        _check(_lib.lt_scorer_create(partition._h, seed, ctypes.byref(h)), "lt_scorer_create")
        self._h = h
        self.P = partition.count

    def add_trace(self, engine, use_next=True, use_keep=True):
        n = _u64()
        _check(_lib.lt_scorer_add_trace(self._h, engine._h, int(use_next), int(use_keep), ctypes.byref(n)),
               "lt_scorer_add_trace")
        return n.value

    def fit(self, epochs=20, batch=32, lr=1e-3, seed=7):
        loss = np.zeros(2, np.float32)
        _check(_lib.lt_scorer_fit(self._h, epochs, batch, lr, seed, loss.ctypes.data), "lt_scorer_fit")
        return loss

    @staticmethod
    def _arrays(samples, P):
        n = len(samples)
        counts = np.ascontiguousarray(np.stack([s["counts"] for s in samples]), dtype=np.uint32)
        resident = np.ascontiguousarray(np.stack([s["resident"] for s in samples]), dtype=np.uint8)
        age = np.ascontiguousarray(np.stack([s["age"] for s in samples]), dtype=np.uint32)
        recent = np.ascontiguousarray(np.stack([s["recent"] for s in samples]), dtype=np.int32)
        cur = np.array([s["cur"] for s in samples], np.int32)
        wt = np.array([s["w_total"] for s in samples], np.float32)
        nxt = np.array([s["next"] for s in samples], np.int32)
        keep = np.ascontiguousarray(np.stack([s["keep"] for s in samples]), dtype=np.float32)
        wa = np.array([s["wa"] for s in samples], np.float32)
        wb = np.array([s["wb"] for s in samples], np.float32)
        assert counts.shape == (n, P)
        return counts, resident, age, recent, cur, wt, nxt, keep, wa, wb

    def loss_grad(self, samples):
        arrs = self._arrays(samples, self.P)
        loss = np.zeros(2, np.float32)
        grad = np.zeros(self.NUM_PARAMS, np.float32)
        _check(_lib.lt_scorer_loss_grad(self._h, len(samples), *[a.ctypes.data for a in arrs], loss.ctypes.data,
                                        grad.ctypes.data), "lt_scorer_loss_grad")
        return loss, grad

    def train(self, samples, epochs=10, batch=32, lr=1e-3, seed=7):
        arrs = self._arrays(samples, self.P)
        loss = np.zeros(2, np.float32)
        _check(_lib.lt_scorer_train(self._h, len(samples), *[a.ctypes.data for a in arrs], epochs, batch, lr, seed,
                                    loss.ctypes.data), "lt_scorer_train")
        return loss

    def _state(self, counts, resident, age, recent):
        out = (np.ascontiguousarray(counts, dtype=np.uint32), np.ascontiguousarray(resident, dtype=np.uint8),
               np.ascontiguousarray(age, dtype=np.uint32), np.ascontiguousarray(recent, dtype=np.int32))
        if not all(a.shape == (self.P,) for a in out[:3]) or out[3].shape != (8,):
            raise ValueError(f"a decision state is counts, resident and age of {self.P} partitions and 8 recent loads")
        return out

    def infer(self, counts, resident, age, recent, cur, w_total):
        c, r, a, rc = self._state(counts, resident, age, recent)
        out = np.zeros((self.P, 2), np.float32)
        _check(_lib.lt_scorer_infer(self._h, c.ctypes.data, r.ctypes.data, a.ctypes.data, rc.ctypes.data, cur,
                                    w_total, out.ctypes.data), "lt_scorer_infer")
        return out

    def infer_noalloc(self, counts, resident, age, recent, cur, w_total, out=None):
        """infer() through the GPU's static scratch (lt_scorer_infer_noalloc);
        `out` (float32 [P, 2]) may be given so that nothing is allocated."""
        c, r, a, rc = self._state(counts, resident, age, recent)
        if out is None:
            out = np.zeros((self.P, 2), np.float32)
        _check(_lib.lt_scorer_infer_noalloc(self._h, c.ctypes.data, r.ctypes.data, a.ctypes.data, rc.ctypes.data,
                                            cur, w_total, out.ctypes.data), "lt_scorer_infer_noalloc")
        return out

    def get_params(self):
        out = np.zeros(self.NUM_PARAMS, np.float32)
        _check(_lib.lt_scorer_get_params(self._h, out.ctypes.data), "lt_scorer_get_params")
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(_lib.lt_scorer_set_params(self._h, p.ctypes.data), "lt_scorer_set_params")

    def save(self, path):
        _check(_lib.lt_scorer_save(self._h, path.encode()), "lt_scorer_save")

    def load(self, path):
        _check(_lib.lt_scorer_load(self._h, path.encode()), "lt_scorer_load")

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.lt_scorer_free(self._h)
        self._h = None

    def __del__(self):
        self.close()
