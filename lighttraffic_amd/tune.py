"""Autotuner of batched PPR: Python view of lt_tune_ppr_batch (no counterpart
in the reference).

The C entry point tries engine configurations on a sample of the queries:
partition size (16/32/64 MB; the cache keeps the same number of bytes), span
(partition | resident), scheduling on the host or on the GPU (the
hipGraph-captured device loop), the number of 4096-walker chunks one walk
launch consumes, (span resident, host scheduling) how many queues one launch
walks (sub-threshold partitions over PCIe, or cached partitions), and whether
partitions are copied raw,
gap-coded and unpacked in HBM (host scheduling), or kept packed in HBM as
16-edge blocks, the latter also on the split sized for block slots of the
partition size (lt_partition_build_blocks, `split` 1).  A small sample walks
few steps per partition load, so it
over-weights the copies: the three fastest configurations, and the fastest of
each copy mode, are timed again on a sample four times larger; the fastest of
those is timed again with 1, 2 and 4 walkers per walk-kernel lane (instead of
the launch-size rule), and the pick is the fastest of the last two stages.
Every configuration produces the same endpoints.
"""
import ctypes

import numpy as np

from . import _check, _lib, _ptr

_SPAN_NAMES = {0: "partition", 1: "resident"}


class TuneTrial(ctypes.Structure):
    """lt_tune_trial (include/lighttraffic.h)."""
    _fields_ = [("stage", ctypes.c_uint32), ("partition_mb", ctypes.c_uint32), ("slots", ctypes.c_uint32),
                ("partitions", ctypes.c_uint32), ("cache_slots", ctypes.c_uint32), ("span", ctypes.c_int32),
                ("device_schedule", ctypes.c_int32), ("slice_chunks", ctypes.c_int32), ("zc_batch", ctypes.c_int32),
                ("pack_partitions", ctypes.c_int32), ("walkers_per_thread", ctypes.c_int32),
                ("split", ctypes.c_int32), ("running_ms", ctypes.c_double), ("h2d_bytes", ctypes.c_uint64),
                ("iterations", ctypes.c_uint64)]

    def as_dict(self):
        d = {name: getattr(self, name) for name, _ in self._fields_}
        d["span"] = _SPAN_NAMES[d["span"]]
        return d


_lib.lt_tune_ppr_batch.restype = ctypes.c_int
_lib.lt_tune_ppr_batch.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float,
                                   ctypes.c_uint64, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
                                   ctypes.POINTER(TuneTrial), ctypes.POINTER(TuneTrial), ctypes.c_uint32,
                                   ctypes.POINTER(ctypes.c_uint32)]


def tune_ppr_batch(graph, sources, counts, prob=0.15, cache_bytes=8 << 30, sample_every=16, sizes_mb=None,
                   log=None):
    """Returns (best, trials) as dicts: trials holds every configuration timed
    on the sample (`stage` 1), the refined ones timed on the larger sample
    (`stage` 2) and stage 2's pick with fixed walkers per lane (`stage` 3);
    best is the fastest of stages 2 and 3."""
    src = np.ascontiguousarray(sources, dtype=np.uint32)
    cnt = np.ascontiguousarray(counts, dtype=np.uint64)
    sizes = None if sizes_mb is None else np.ascontiguousarray(sizes_mb, dtype=np.uint32)
    cap = 256
    buf = (TuneTrial * cap)()
    best = TuneTrial()
    n = ctypes.c_uint32()
    _check(_lib.lt_tune_ppr_batch(graph._h, len(src), src.ctypes.data, cnt.ctypes.data, prob, cache_bytes,
                                  _ptr(sizes), 0 if sizes is None else len(sizes), sample_every, ctypes.byref(best),
                                  buf, cap, ctypes.byref(n)), "lt_tune_ppr_batch")
    trials = [buf[i].as_dict() for i in range(min(n.value, cap))]
    if log:
        for t in trials:
            log(f"tune {t}")
    return best.as_dict(), trials
