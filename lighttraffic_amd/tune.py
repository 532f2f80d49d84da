"""Autotuner for batched PPR (no counterpart in the reference).

Tries engine configurations on a sample of the queries and returns the
fastest: partition size (16/32/64 MB, the cache keeps the same number of
bytes), span (partition | resident), scheduling on the host or on the GPU
(hipGraph-captured device loop), the number of 4096-walker chunks one walk
launch consumes, (span resident, host scheduling) how many sub-threshold
partitions one zero-copy launch walks, and (host scheduling) whether partitions
are copied raw, gap-coded and unpacked in HBM, or kept packed in HBM as
16-edge blocks (the same cache bytes then hold more partitions).

A small sample walks few steps per partition load, so it over-weights the
copies.  The `refine` fastest configurations, and the fastest of each copy
mode, are therefore timed again on a sample `refine_factor` times larger, and
the pick is the fastest of those.
Every configuration produces the same endpoints, so the choice only affects
speed and PCIe bytes.
"""
import itertools
import time

import numpy as np

from . import Engine, Partition

_KEYS = ("span", "device_schedule", "slice_chunks", "zc_batch", "pack_partitions")


def _time(graph, part, slots, src, cnt, prob, cfg):
    """Statistics and wall time of a warm second job of `cfg` on the query sample."""
    eng = Engine(graph, part, num_device_partitions=slots, **cfg)
    eng.ppr_batch(src, cnt, prob=prob, endpoints=False)  # warm the cache
    t0 = time.perf_counter()
    eng.ppr_batch(src, cnt, prob=prob, endpoints=False)
    wall = time.perf_counter() - t0
    st = eng.stats()
    eng.close()
    return st, wall


def tune_ppr_batch(graph, sources, counts, prob=0.15, cache_bytes=8 << 30, sample_every=16,
                   sizes_mb=(16, 32, 64), spans=("partition", "resident"), device_schedule=(0, 1),
                   slices=(2048, 8192), zc_batches=(1, 16), packs=(0, 1, 2), refine=3, refine_factor=4,
                   log=None):
    """Returns (best, trials): trials holds one dict per configuration with its
    time on the sample (`stage` 1), then the refined configurations with their
    time on the larger sample (`stage` 2); best is the fastest of the last
    stage run."""
    def sample(every):
        sel = np.arange(0, len(sources), max(1, every))
        return np.asarray(sources)[sel], np.asarray(counts)[sel]

    s_src, s_cnt = sample(sample_every)
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
            cfg = dict(span=span, device_schedule=dev, slice_chunks=sl, zc_batch=zb, pack_partitions=pk)
            st, wall = _time(graph, part, slots, s_src, s_cnt, prob, cfg)
            rec = {"stage": 1, "partition_mb": mb, "slots": slots, "partitions": part.count, **cfg,
                   "cache_slots": st["cache_slots"], "running_ms": st["running_ms"],
                   "wall_ms": wall * 1e3, "h2d_bytes": st["h2d_bytes_explicit"] + st["h2d_bytes_zerocopy"],
                   "iterations": st["iterations"]}
            trials.append(rec)
            if log:
                log(f"tune {rec}")
        graph.drop_packed()  # the next partition size needs images of its own
        part.close()
    last = trials
    if refine > 1 and refine_factor > 1 and sample_every > 1:
        top = sorted(trials, key=lambda r: r["running_ms"])[:refine]
        for pk in packs:  # the copy modes trade bytes for decode work: compare each on the larger sample
            mode = [t for t in trials if t["pack_partitions"] == pk]
            if mode:
                fastest = min(mode, key=lambda r: r["running_ms"])
                if fastest not in top:
                    top.append(fastest)
        r_src, r_cnt = sample(sample_every // refine_factor)
        last = []
        for mb in sorted({t["partition_mb"] for t in top}):
            part = Partition.build(graph, mb << 20)
            for t in (t for t in top if t["partition_mb"] == mb):
                cfg = {k: t[k] for k in _KEYS}
                st, wall = _time(graph, part, t["slots"], r_src, r_cnt, prob, cfg)
                rec = dict(t, stage=2, running_ms=st["running_ms"], wall_ms=wall * 1e3,
                           h2d_bytes=st["h2d_bytes_explicit"] + st["h2d_bytes_zerocopy"],
                           iterations=st["iterations"])
                last.append(rec)
                if log:
                    log(f"tune refine {rec}")
            graph.drop_packed()
            part.close()
        trials = trials + last
    best = min(last, key=lambda r: r["running_ms"])
    return best, trials
