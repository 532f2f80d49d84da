"""Autotuner for batched PPR (no counterpart in the reference).

Tries engine configurations on a sample of the queries and returns the
fastest: partition size (16/32/64 MB, the cache keeps the same number of
bytes), span (partition | resident), scheduling on the host or on the GPU
(hipGraph-captured device loop), the number of 4096-walker chunks one walk
launch consumes, (span resident, host scheduling) how many sub-threshold
partitions one zero-copy launch walks, and (host scheduling) whether partitions
are copied raw, gap-coded and unpacked in HBM, or kept packed in HBM as
16-edge blocks (the same cache bytes then hold more partitions).  Every configuration produces the same endpoints, so the
choice only affects speed and PCIe bytes.
"""
import itertools
import time

import numpy as np

from . import Engine, Partition


def tune_ppr_batch(graph, sources, counts, prob=0.15, cache_bytes=8 << 30, sample_every=16,
                   sizes_mb=(16, 32, 64), spans=("partition", "resident"), device_schedule=(0, 1),
                   slices=(2048, 8192), zc_batches=(1, 16), packs=(0, 1, 2), log=None):
    """Returns (best, trials): best is a dict of Partition/Engine settings, trials
    one dict per configuration with its time on the sample."""
    sel = np.arange(0, len(sources), sample_every)
    s_src, s_cnt = np.asarray(sources)[sel], np.asarray(counts)[sel]
    trials = []
    for mb in sizes_mb:
        part = Partition.build(graph, mb << 20)
        slots = max(1, min(part.count, cache_bytes // (mb << 20)))
        for span, dev, sl, zb, pk in itertools.product(spans, device_schedule, slices, zc_batches, packs):
            if dev and sl != slices[0]:
                continue  # the device loop consumes whole queues; slices do not apply
            if zb != zc_batches[0] and (dev or span != "resident"):
                continue  # zero-copy batches need host scheduling and the resident span
            if pk != packs[0] and dev:
                continue  # the device loop copies partitions raw
            eng = Engine(graph, part, num_device_partitions=slots, span=span, device_schedule=dev, slice_chunks=sl,
                         zc_batch=zb, pack_partitions=pk)
            eng.ppr_batch(s_src, s_cnt, prob=prob, endpoints=False)  # warm the cache
            t0 = time.perf_counter()
            eng.ppr_batch(s_src, s_cnt, prob=prob, endpoints=False)
            wall = time.perf_counter() - t0
            st = eng.stats()
            eng.close()
            rec = {"partition_mb": mb, "slots": slots, "partitions": part.count, "span": span,
                   "device_schedule": dev, "slice_chunks": sl, "zc_batch": zb, "pack_partitions": pk,
                   "cache_slots": st["cache_slots"], "running_ms": st["running_ms"],
                   "wall_ms": wall * 1e3, "h2d_bytes": st["h2d_bytes_explicit"] + st["h2d_bytes_zerocopy"],
                   "iterations": st["iterations"]}
            trials.append(rec)
            if log:
                log(f"tune {rec}")
        graph.drop_packed()  # the next partition size needs images of its own
        part.close()
    best = min(trials, key=lambda r: r["running_ms"])
    return best, trials
