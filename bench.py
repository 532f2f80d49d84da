"""Benchmark of the MI355X random-walk engine (one JSON line on stdout).

A step is one complete walk job over one synthetic R-MAT graph: every walker
created, walked to its end under the partition scheduler, results reduced.
Timed region per step = the scheduler loop ("Running Time" of
src/Scheduler.h:425-441): graph generation, walker creation and the copy of
results back to the host are outside it, as in the reference.  Each measured
configuration also reports the median of its steps and the end-to-end wall
time of a job (walker creation and result copy included).  The pinned
host-to-HBM copy rate is measured once and reported beside the link's
theoretical rate.

Workloads (R-MAT a,b,c,d = .57/.19/.19/.05, seeded, deduped, symmetrised,
self loops and isolated vertices removed):
  c5  (default) scale 26 ef 27, 32 MB partitions, 8 GB partition cache
      (256 slots), pagerank w=2|V| l=10: plain scheduler vs learned scheduler
  c1  scale 16 ef 16, one 32 MB partition loaded into HBM (zero-copy threshold 0),
      pagerank w=2|V| l=10
  c2  scale 22 ef 16, 32 MB partitions all HBM-resident, genericwalk w=2|V| l=10,
      walks continue through every partition (the reference's span is reported beside)
  c3  scale 25 ef 44, 4 GB partition cache, 2^26 walkers at once, ppr from vertex 0;
      value: span resident, zero-copy batches of 16, partitions kept packed in
      HBM (the reference's rules and its figure-14 ablations are reported beside)
  c4  same graph, 4,096 PPR queries (sources proportional to degree, walk counts
      log-uniform in [1e3, 1e6]), 8 GB cache; configuration chosen by the
      autotuner (partition size 16/32/64 MB, span, host or device scheduling,
      zero-copy batches, raw / gap-coded / block copies)
For c5 the learned scorer is trained here, on the GPU, from the traces of the
plain scheduler's warm-up runs; `value` is the learned scheduler's steps/s with
partitions kept packed in HBM (16-edge frame-of-reference blocks, copied as
they are and decoded by the walk, so the 8 GB cache holds more partitions).
Beside it: the plain scheduler as the reference runs it (raw copies), the plain
scheduler with the same block copies and with gap-coded copies unpacked in HBM,
and the learned scheduler with gap-coded and with raw copies.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "walk steps/s & H2D bytes/step: learned vs plain scheduler, 1 GPU"
HBM_PEAK_GBS = 8000.0
# Philox key of the walks the scorer is trained on (the timed walks use the
# engine's default key, 0x4C54524B): training and timing see different jobs
TRAIN_SEED = 0x4C54524C
# C5's learned configuration splits the graph for block slots of this size
# (lt_partition_build_blocks) instead of the reference's 32 MB raw split
BLOCK_SLOT = 32 << 20

WORKLOADS = {
    "c1": dict(scale=16, ef=16, mb=32, app="pagerank", slots=0, l=10, zc=0),
    "c2": dict(scale=22, ef=16, mb=32, app="genericwalk", slots=0, l=10, span="resident",
               eng=dict(resident_batch=16, walkers_per_thread=2)),
    # c3's value: the prefetching loop without a network (largest queues needing
    # host data copied ahead, plain victims) on the graph split for 32 MiB
    # block slots (lt_partition_build_blocks) in the same 4 GB, zero-copy
    # threshold slot/128, with the whole rowptr in HBM (170 MB of the cache) so
    # that a zero-copied move reads only its column over PCIe; the variants
    # use the reference's 32 MB split and its S/256 threshold
    "c3": dict(scale=25, ef=44, mb=32, app="ppr", slots=128, l=0, wdev=1 << 26, zc_div=256, blocks_mb=32,
               blocks_zc_div=128,
               eng=dict(span="resident", zc_batch=16, resident_batch=16, pack_partitions=2, policy="learned",
                        prefetch_plain=1, walkers_per_thread=1, rowptr_resident=1)),
    "c4": dict(scale=25, ef=44, mb=32, app="ppr_batch", slots=256, l=0, nq=4096),
    "c5": dict(scale=26, ef=27, mb=32, app="pagerank", slots=256, l=10),
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes(st):
    """rowptr pair 16 B + col 4 B per move, a u32 read-modify-write per visit,
    one 16-byte walker record read and written per queue residency."""
    return 20 * st["moves"] + 8 * st["visits"] + 32 * st["residencies"]


QUERIES = {}


def run_job(eng, wl, w):
    if wl["app"] == "ppr_batch":
        src, cnt = QUERIES["table"]
        eng.ppr_batch(src, cnt, prob=0.15, endpoints=False)
    elif wl["app"] == "pagerank":
        eng.pagerank(w, wl["l"])
    elif wl["app"] == "genericwalk":
        eng.genericwalk(w, wl["l"], endpoints=True)
    else:
        eng.ppr(w, source=0, prob=0.15)
    return eng.stats()


def measure(eng, wl, w, warmup, steps, label, timed_engine=None):
    """Warm-up jobs, then `steps` timed jobs on an engine without event
    timers; the walk-kernel time for the roofline comes from one more job on
    `timed_engine` (same configuration, HIP event timers on), so the timers'
    cost stays out of the steps/s figures."""
    for i in range(warmup):
        st = run_job(eng, wl, w)
        log(f"{label} warmup {i}: {st['running_ms']:.1f} ms")
    stats, walls = [], []
    for i in range(steps):
        t0 = time.perf_counter()
        stats.append(run_job(eng, wl, w))
        walls.append((time.perf_counter() - t0) * 1e3)
        s = stats[-1]
        log(f"{label} step {i}: {s['running_ms']:.1f} ms, loads {s['graph_loading_calls']}, "
            f"H2D {(s['h2d_bytes_explicit'] + s['h2d_bytes_zerocopy']) / 1e9:.2f} GB")
    ms = float(np.mean([s["running_ms"] for s in stats]))
    nominal = stats[0]["nominal_steps"]
    keys = ("moves", "visits", "residencies", "kernel_launches", "h2d_bytes_explicit", "h2d_bytes_zerocopy",
            "graph_loading_calls", "zerocopy_calls", "iterations", "prefetch_issued", "prefetch_hits",
            "h2d_bytes_unpacked", "packed_image_bytes", "packed_col_bytes")
    tot = {k: sum(s[k] for s in stats) for k in keys}
    h2d = (tot["h2d_bytes_explicit"] + tot["h2d_bytes_zerocopy"]) / len(stats)
    res = {"ms_per_step": ms, "steps_per_s": nominal / (ms * 1e-3), "h2d_bytes_per_step": h2d / nominal,
           "median_ms_per_step": float(np.median([s["running_ms"] for s in stats])),
           "e2e_ms_per_step": float(np.mean(walls)),
           "h2d_GB_per_job": h2d / 1e9, "pcie_GBps": h2d / (ms * 1e-3) / 1e9,
           "per_job": {k: v / len(stats) for k, v in tot.items()}, "nominal": nominal}
    if timed_engine is not None:
        tst = run_job(timed_engine, wl, w)
        res.update(_timed(tst))
    return res


def _timed(tst):
    return {"kernel_ms": tst["walk_kernel_ms"], "launches": tst["kernel_launches"],
            "alg_bytes": algorithmic_bytes(tst), "timed_job_ms": tst["running_ms"],
            "zc_launches": tst["zerocopy_launches"], "zc_kernel_ms": tst["walk_kernel_zc_ms"]}


def timed_walk(eng, wl, w):
    """Walk-kernel time of one job for the roofline: a warm-up job, then one
    job on an engine with HIP event timers around every walk launch."""
    run_job(eng, wl, w)
    return _timed(run_job(eng, wl, w))


def measured_traffic(workload, span, alg_bytes_per_launch):
    """HBM bytes per walk-kernel launch from the committed rocprofv3 counter
    summary of this workload (FETCH_SIZE + WRITE_SIZE, calibrated on a copy
    kernel of known size), as the ratio traffic / algorithmic bytes measured
    there times this run's algorithmic bytes per launch; None if not profiled."""
    path = os.path.join(ROOT, "profiles", f"walk_kernel_{workload}_{span}_pmc.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d["traffic_per_algorithmic_byte"] * alg_bytes_per_launch


def roofline(m, workload=None, span=None):
    achieved = m["alg_bytes"] / (m["kernel_ms"] * 1e-3) / 1e9 if m["kernel_ms"] > 0 else 0.0
    per_launch = m["alg_bytes"] / max(m["launches"], 1)
    traffic = measured_traffic(workload, span, per_launch)
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "traffic_source": None if traffic is None else
            (f"derived, not measured in this run: the FETCH_SIZE + WRITE_SIZE / algorithmic-byte ratio of "
             f"profiles/walk_kernel_{workload}_{span}_pmc.json times this run's algorithmic bytes per launch"),
            "kernel": "walk_kernel",
            "launches": m["launches"], "avg_launch_us": m["kernel_ms"] * 1e3 / max(m["launches"], 1),
            "algorithmic_bytes_per_launch": m["alg_bytes"] / max(m["launches"], 1),
            "zerocopy_launches": m.get("zc_launches"), "zerocopy_kernel_ms": m.get("zc_kernel_ms")}


def roofline_pcie(m, peak_gbps):
    """Capped-cache workloads are bound by the host link: H2D bytes (explicit
    partition copies + zero-copied reads + spilled walkers) per second of the
    timed region against the measured pinned copy rate, and the steps/s that
    rate allows at this run's H2D bytes per step."""
    achieved = m["pcie_GBps"]
    return {"bound": "pcie", "achieved": achieved, "peak": peak_gbps, "unit": "GB/s", "frac": achieved / peak_gbps,
            "peak_kind": "measured pinned hipMemcpyAsync rate", "theoretical": 64.0,
            "h2d_bytes_per_step": m["h2d_bytes_per_step"],
            "bound_steps_per_s": peak_gbps * 1e9 / m["h2d_bytes_per_step"] if m["h2d_bytes_per_step"] > 0 else None}


def belady_bound(eng, part):
    """The engine's last job replayed with Belady's optimal eviction: the
    partitions it walked from HBM (trace records of kind 0, 2 and 3), in
    order, through the same number of cache slots, starting from the job's
    first residency.  Its loads are the fewest any victim rule needs for this
    processing order; the H2D floor scales the job's explicit bytes by them."""
    import lighttraffic_amd as lt
    tr = eng.trace()
    _, resident, _, _ = eng.trace_state()
    st = eng.stats()
    seq = [r[0] for r in tr if r[1] in (0, 2, 3)]
    loads = st["graph_loading_calls"]
    best, _, _ = lt.sched_belady(seq, part.count, st["cache_slots"], init_resident=resident[0])
    explicit = st["h2d_bytes_explicit"] * (best / loads if loads else 1.0)
    return {"demands": len(seq), "loads": loads, "loads_belady": best,
            "load_saving_bound": 1.0 - best / loads if loads else 0.0,
            "h2d_bytes_per_step": (st["h2d_bytes_explicit"] + st["h2d_bytes_zerocopy"]) / st["nominal_steps"],
            "h2d_bytes_per_step_floor": (explicit + st["h2d_bytes_zerocopy"]) / st["nominal_steps"]}


def cpu_baseline(g, wl, w, budget_s=20.0):
    """The oracle (OpenMP, whole graph in DRAM) on a bounded sample of the
    same workload; steps/s scaled to the sample."""
    from oracle import ffi as oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    rowptr, col = g.rowptr(), g.col()
    n = min(w, 200000)
    while True:
        t0 = time.perf_counter()
        if wl["app"] == "ppr_batch":
            src, cnt = QUERIES["table"]
            k = int(np.searchsorted(np.cumsum(cnt), n)) + 1
            oracle.ppr_batch(rowptr, col, src[:k], cnt[:k], threads=threads)
            steps = float(cnt[:k].sum()) / 0.15
            n = int(cnt[:k].sum())
        elif wl["app"] == "pagerank":
            oracle.pagerank(rowptr, col, w, wl["l"], id_begin=0, id_end=n, threads=threads)
            steps = n * wl["l"]
        elif wl["app"] == "genericwalk":
            oracle.genericwalk(rowptr, col, w, wl["l"], id_begin=0, id_end=n, threads=threads)
            steps = n * wl["l"]
        else:
            oracle.ppr(rowptr, col, w, source=0, id_begin=0, id_end=n, threads=threads)
            steps = n / 0.15
        dt = time.perf_counter() - t0
        if dt > budget_s / 4 or n >= w:
            break
        n = min(w, int(n * max(2.0, (budget_s / 2) / max(dt, 1e-3))))
    if wl["app"] == "ppr_batch":
        w = int(QUERIES["table"][1].sum())
    return {"value": steps / dt, "unit": "walk steps/s", "cores": threads, "kind": "port",
            "sample": f"oracle {wl['app']} over walker ids [0, {n}) of {w} on the same graph, "
                      f"{dt:.2f} s on {threads} threads ({_cpu_model()})"}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown cpu"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c5", choices=sorted(WORKLOADS))
    ap.add_argument("--span", default=None, choices=["partition", "resident"],
                    help="single-policy workloads only (default partition)")
    ap.add_argument("--epochs", type=int, default=20, help="scorer training epochs (c5)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--c4-config", default=None, metavar="JSON",
                    help="c4: skip the autotuner and run this configuration (a JSON object with the fields of the "
                         "autotuner's pick: partition_mb, slots, span, device_schedule, slice_chunks, zc_batch, "
                         "pack_partitions, walkers_per_thread, split)")
    ap.add_argument("--no-variants", action="store_true",
                    help="c2, c3: measure the value's configuration only, without the other configurations "
                         "the line reports beside it (for profiling runs)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run copy_probe_kernel over MIB MiB, twice (a kernel of known traffic for "
                         "scaling rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU sharding is a design note only",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    # Before the HIP runtime starts: at least eight hardware queues, so that a
    # packed copy's DMA stream gets a queue of its own (the engine checks this
    # value; HIP's default is 4).
    if int(os.environ.get("GPU_MAX_HW_QUEUES", "0") or 0) < 8:
        os.environ["GPU_MAX_HW_QUEUES"] = "8"
    import lighttraffic_amd as lt

    wl = WORKLOADS[args.workload]
    info = lt.device_info()
    if args.copy_calibration:
        gbps = lt.device_copy_bandwidth(args.copy_calibration << 20, reps=1)
        log(f"copy probe {args.copy_calibration} MiB: {gbps:.0f} GB/s (read + write)")
    h2d_peak = lt.device_h2d_bandwidth(256 << 20, reps=4)
    log(f"pinned H2D copy: {h2d_peak:.1f} GB/s")
    t0 = time.perf_counter()
    g = lt.Graph.rmat(wl["scale"], wl["ef"])
    log(f"graph s{wl['scale']} ef{wl['ef']}: V={g.num_node} E={g.num_edge} in {time.perf_counter() - t0:.1f} s")
    part = lt.Partition.build(g, wl["mb"] << 20)
    w = 2 * g.num_node
    base = dict(num_device_partitions=wl["slots"])
    if "wdev" in wl:
        base["num_device_walkers"] = wl["wdev"]
    if "zc" in wl:  # c1: load the one partition (its 2|V| walkers are below S/256) so it stays cached
        base["zerocopy_threshold"] = wl["zc"]
    if "zc_div" in wl:
        base["zerocopy_threshold"] = part.bytes // wl["zc_div"]
    desc = (f"{args.workload}: R-MAT scale {wl['scale']} ef {wl['ef']} (V={g.num_node}, E={g.num_edge}), "
            f"{wl['app']} w=2|V|={w}" + (f" l={wl['l']}" if wl['l'] else " p=0.15") +
            f", {wl['mb']} MB partitions (P={part.count}), cache " +
            ("all partitions" if not wl["slots"] else f"{wl['slots']} slots = {wl['slots'] * wl['mb'] // 1024} GB"))
    out = {"metric": METRIC, "unit": "walk steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
           "data": "synthetic R-MAT (a,b,c,d=.57/.19/.19/.05), generated on the GPU, Philox walks",
           "pcie": {"h2d_measured_peak_GBps": h2d_peak, "h2d_theoretical_GBps": 64.0,
                    "note": "pinned hipMemcpyAsync of 256 MiB, 4 times; PCIe 5.0 x16 theoretical rate"}}

    if args.workload == "c5":
        # One engine at a time (plus the scorer's stream): HIP maps streams onto
        # GPU_MAX_HW_QUEUES hardware queues (4 by default), and an engine whose
        # copy and compute streams share one loses its copy/compute overlap.
        plain = {}
        keep = ("steps_per_s", "ms_per_step", "median_ms_per_step", "e2e_ms_per_step", "h2d_bytes_per_step",
                "h2d_GB_per_job", "pcie_GBps")
        # The reference's 32 MB split (EqualSizeStrategy): its rules with raw,
        # gap-coded (1) and block (2) copies, span partition and resident
        for span in ("partition", "resident"):
            eng = lt.Engine(g, part, span=span, **base)
            plain[span] = measure(eng, wl, w, args.warmup, args.steps, f"plain/{span}")
            if span == "resident":
                ref = eng.pagerank(w, wl["l"])
            eng.close()
            for mode, tag in ((1, "_packed"), (2, "_blocks")):
                eng = lt.Engine(g, part, span=span, pack_partitions=mode, **base)
                plain[span + tag] = measure(eng, wl, w, args.warmup, args.steps, f"plain/{span}{tag}")
                eng.close()
        base["walkers_per_thread"] = 1  # from here on: 1 walker per lane measured 1-2 % faster
        # the same rules with zero-copy threshold S/192 and zero-copy and resident
        # batches of 16, on that split
        for mode, tag in ((1, "_packed"), (2, "_blocks_32MB_split")):
            eng = lt.Engine(g, part, span="resident", zerocopy_threshold=part.bytes // 192, zc_batch=16,
                            resident_batch=16, pack_partitions=mode, **base)
            plain["tuned" + tag] = measure(eng, wl, w, args.warmup, args.steps,
                                           f"plain/resident S/192 batches 16{tag}")
            eng.close()
        g.drop_packed()  # the 32 MB split's images; the balanced split gets its own
        # The learned configuration's table: partitions split so that each one's
        # rowptr slice + block image fills a 32 MiB block slot
        # (lt_partition_build_blocks); the 8 GB cache then holds 0.87 of the
        # graph's partitions instead of 0.72 with the 32 MB split, whose block
        # slots are sized by its largest image.
        t1 = time.perf_counter()
        pblk = lt.Partition.build_blocks(g, BLOCK_SLOT)
        split_s = time.perf_counter() - t1
        bbase = dict(base, num_device_partitions=(wl["slots"] * (wl["mb"] << 20)) // pblk.bytes)
        log(f"balanced split: {pblk.count} partitions of <= {BLOCK_SLOT >> 20} MiB block slots in {split_s:.1f} s")
        # The learned configuration's settings (a sweep over prefetch depth 2/3
        # and thresholds slot/128 and slot/192 picked them)
        zt, depth = BLOCK_SLOT // 128, 2
        common = dict(span="resident", zerocopy_threshold=zt, zc_batch=16, resident_batch=16, pack_partitions=2)
        belady = {}
        # plain rules with all of the learned configuration's non-learned settings
        eng = lt.Engine(g, pblk, record_trace=1, **common, **bbase)
        plain["tuned_blocks"] = measure(eng, wl, w, args.warmup, args.steps, "plain/balanced blocks slot/128")
        belady["plain_same_settings"] = belady_bound(eng, pblk)
        eng.close()
        # the learned loop without its network: prefetch the largest queues
        # that need host data, evict by the plain rule (what the network adds)
        eng = lt.Engine(g, pblk, policy="learned", prefetch_plain=1, prefetch_depth=depth, **common, **bbase)
        plain["prefetch_rule_blocks"] = measure(eng, wl, w, args.warmup, args.steps, "plain-rule prefetch blocks")
        eng.close()
        # Training traces: the plain rules with the learned configuration's other
        # settings (threshold slot/192), on walks of another Philox key (held
        # out from the timed jobs), warm jobs only (the first fills an empty
        # cache and evicts nothing).  Labels: the next partition needing host
        # data (output 0) and Belady's next-use quantity (output 1, which
        # picks victims; prefetch targets follow the plain rule).
        t1 = time.perf_counter()
        sc = lt.Scorer(pblk, seed=1)
        eng = lt.Engine(g, pblk, record_trace=1, seed=TRAIN_SEED, **dict(common, zerocopy_threshold=BLOCK_SLOT // 192),
                        **bbase)
        n_train = 0
        for i in range(4):
            run_job(eng, wl, w)
            if i > 0:
                n_train += sc.add_trace(eng, use_next=True, use_keep=2)
        eng.close()
        prep_s = time.perf_counter() - t1
        t1 = time.perf_counter()
        loss = sc.fit(epochs=args.epochs, batch=32, lr=1e-3)
        train_s = time.perf_counter() - t1
        log(f"scorer trained on {n_train} decisions in {train_s:.1f} s (traces {prep_s:.1f} s), "
            f"losses {loss.tolist()}")
        lcfg = dict(common, policy="learned", prefetch_plain=2, prefetch_depth=depth)
        # what the network's victim choice adds: the same loop evicting the
        # least recently walked candidate (shortlist 1), and the network
        # choosing among the 4 least recently walked (shortlist 4)
        ablate = {}
        for k, tag in ((1, "lru_only"), (4, "shortlist4")):
            eng = lt.Engine(g, pblk, evict_shortlist=k, **lcfg, **bbase)
            eng.set_scorer(sc)
            ablate[tag] = measure(eng, wl, w, args.warmup, args.steps, f"learned blocks, shortlist {k}")
            eng.close()
        # the network also choosing the next partition among near-ties of the
        # greedy choice (queues within 50 % of it)
        eng = lt.Engine(g, pblk, learned_next=50, **lcfg, **bbase)
        eng.set_scorer(sc)
        ablate["next50"] = measure(eng, wl, w, args.warmup, args.steps, "learned blocks, next by the network")
        ablate["next50"]["next_overrides_last_job"] = eng.stats()["next_overrides"]
        eng.close()
        eng = lt.Engine(g, pblk, record_trace=1, **lcfg, **bbase)
        eng.set_scorer(sc)
        learned = measure(eng, wl, w, args.warmup, args.steps, "learned/resident blocks")
        belady["learned"] = belady_bound(eng, pblk)
        same = bool(np.array_equal(eng.pagerank(w, wl["l"]), ref))
        slots = eng.stats()["cache_slots"]
        eng.close()
        teng = lt.Engine(g, pblk, event_timers=1, **lcfg, **bbase)
        teng.set_scorer(sc)
        learned.update(timed_walk(teng, wl, w))
        teng.close()
        ref_plain = plain["partition"]
        same_copy = {k: v for k, v in plain.items() if "blocks" in k}

        def vs(m):
            return {"speedup": learned["steps_per_s"] / m["steps_per_s"],
                    "h2d_bytes_ratio": learned["h2d_bytes_per_step"] / m["h2d_bytes_per_step"]}

        out.update({
            "value": learned["steps_per_s"],
            "ms_per_step": learned["ms_per_step"],
            "median_ms_per_step": learned["median_ms_per_step"],
            "e2e_ms_per_step": learned["e2e_ms_per_step"],
            "config": {"workload": desc, "scheduler": "learned (victims: fewest queued walkers, then the BF16 "
                                                      "MLP's lowest keep score; "
                                                      f"prefetch depth {depth} of the largest queues needing host "
                                                      "data; span resident, zero-copy threshold slot/128, "
                                                      "zero-copy and resident batches of 16)",
                       "partitions": (f"{pblk.count} partitions sized for {BLOCK_SLOT >> 20} MiB block slots "
                                      f"(lt_partition_build_blocks), kept packed in HBM as 16-edge "
                                      f"frame-of-reference blocks, {slots} cache slots in the 8 GB; the plain "
                                      f"variants named *_32MB_split and span_* use the reference's 32 MB split "
                                      f"(P={part.count})"),
                       "device": info["name"]},
            "h2d_bytes_per_step": learned["h2d_bytes_per_step"],
            "plain": {"span_partition": {k: ref_plain[k] for k in keep},
                      "span_partition_packed": {k: plain["partition_packed"][k] for k in keep},
                      "span_partition_blocks": {k: plain["partition_blocks"][k] for k in keep},
                      "span_resident": {k: plain["resident"][k] for k in keep},
                      "span_resident_packed": {k: plain["resident_packed"][k] for k in keep},
                      "span_resident_blocks": {k: plain["resident_blocks"][k] for k in keep},
                      "span_resident_S192_batch16_packed": {k: plain["tuned_packed"][k] for k in keep},
                      "span_resident_S192_batch16_blocks_32MB_split": {
                          k: plain["tuned_blocks_32MB_split"][k] for k in keep},
                      "balanced_blocks_slot128_batch16": {k: plain["tuned_blocks"][k] for k in keep},
                      "plain_rule_prefetch_blocks": {k: plain["prefetch_rule_blocks"][k] for k in keep}},
            "learned": {"blocks": {k: learned[k] for k in keep},
                        "blocks_lru_only": {k: ablate["lru_only"][k] for k in keep},
                        "blocks_shortlist4": {k: ablate["shortlist4"][k] for k in keep},
                        "blocks_next50": {k: ablate["next50"][k] for k in keep + ("next_overrides_last_job",)}},
            "learned_vs_plain": {
                # the headline pair: the plain rules with every non-learned setting of the
                # learned configuration (split, span, zero-copy threshold and batches, block copies)
                "like_for_like": dict(vs(plain["tuned_blocks"]), plain="balanced_blocks_slot128_batch16",
                                      target={"speedup": 1.3, "h2d_bytes_ratio": 0.75}),
                "vs_plain_rule_prefetch": dict(vs(plain["prefetch_rule_blocks"]), note=(
                    "the same prefetching loop with the same targets and the plain victim rule")),
                "vs_lru_only": dict(vs(ablate["lru_only"]), note=(
                    "the same loop evicting the least recently walked fewest-walker candidate: what the "
                    "network's victim choice adds")),
                "vs_shortlist4": dict(vs(ablate["shortlist4"]), note=(
                    "the network choosing among the 4 least recently walked fewest-walker candidates")),
                "vs_next50": dict(vs(ablate["next50"]), note=(
                    "the network also choosing the next partition (learned_next 50): when the greedy "
                    "choice needs host data, the highest 'next' score among queues within 50 % of it")),
                "vs_reference_plain": vs(ref_plain),
                "vs_plain_same_copies": {k: vs(v) for k, v in same_copy.items()},
                "balanced_split_gain_plain": {
                    "speedup": plain["tuned_blocks"]["steps_per_s"] / plain["tuned_blocks_32MB_split"]["steps_per_s"],
                    "h2d_bytes_ratio": (plain["tuned_blocks"]["h2d_bytes_per_step"]
                                        / plain["tuned_blocks_32MB_split"]["h2d_bytes_per_step"]),
                    "note": "the plain rules on the balanced split against the same rules on the 32 MB split "
                            "(threshold slot/128 against S/192)"},
                "belady": belady,
                "visit_counts_identical": same,
                "prefetch_issued_per_job": learned["per_job"]["prefetch_issued"],
                "prefetch_hits_per_job": learned["per_job"]["prefetch_hits"],
                "scorer_training_s": train_s, "training_decisions": n_train,
                "training_walks": f"3 warm pagerank jobs of the plain rules on the balanced split at slot/192, "
                                  f"Philox key {TRAIN_SEED:#x} (timed jobs: 0x4c54524b)",
                "labels": "next partition needing host data (output 0), Belady next-use (output 1, use_keep=2)"},
            "roofline": roofline(learned, "c5", "resident"),
            "roofline_pcie": roofline_pcie(learned, h2d_peak),
            "stats_per_step": learned["per_job"],
        })
    elif args.workload == "c4":
        from lighttraffic_amd.tune import tune_ppr_batch
        src, cnt = g.ppr_queries(wl["nq"])
        QUERIES["table"] = (src, cnt)
        t1 = time.perf_counter()
        if args.c4_config:  # a fixed configuration, e.g. to profile the walk without the tuner's trials
            best, trials = dict(json.loads(args.c4_config), stage=0, cache_slots=0), []
        else:
            best, trials = tune_ppr_batch(g, src, cnt, cache_bytes=wl["slots"] * (wl["mb"] << 20), log=log)
        tune_s = time.perf_counter() - t1
        log(f"autotuner picked {best} in {tune_s:.1f} s")
        part = (lt.Partition.build_blocks if best.get("split") else lt.Partition.build)(g, best["partition_mb"] << 20)
        cfg = dict(num_device_partitions=best["slots"], span=best["span"], device_schedule=best["device_schedule"],
                   slice_chunks=best["slice_chunks"], zc_batch=best["zc_batch"], resident_batch=best["zc_batch"],
                   pack_partitions=best["pack_partitions"], walkers_per_thread=best["walkers_per_thread"])
        eng = lt.Engine(g, part, **cfg)
        # walk-kernel time for the roofline: the same partitioning, host-scheduled, event timers on
        teng = lt.Engine(g, part, **dict(cfg, device_schedule=0), event_timers=1)
        run_job(teng, wl, w)
        m = measure(eng, wl, w, args.warmup, args.steps, args.workload, timed_engine=teng)
        desc = (f"c4: R-MAT scale {wl['scale']} ef {wl['ef']} (V={g.num_node}, E={g.num_edge}), batched ppr "
                f"{wl['nq']} queries (sources ~ degree, counts log-uniform [1e3,1e6), {int(cnt.sum())} walkers), "
                f"p=0.15, 8 GB cache; {'autotuned' if trials else 'fixed configuration'}: "
                f"{best['partition_mb']} MB " + ("block slots, the graph split for them" if best.get("split")
                                                   else "partitions") + f" (P={part.count}), "
                f"{eng.stats()['cache_slots']} cache slots, span {best['span']}, "
                f"{'device' if best['device_schedule'] else 'host'} scheduling, slice {best['slice_chunks']}, "
                f"zero-copy and resident batches of {best['zc_batch']}, "
                f"{('raw', 'gap-coded', 'block')[best['pack_partitions']]} copies, walkers per lane "
                f"{best['walkers_per_thread'] or 'by launch size'}")
        out.update({"value": m["steps_per_s"], "ms_per_step": m["ms_per_step"],
                    "median_ms_per_step": m["median_ms_per_step"], "e2e_ms_per_step": m["e2e_ms_per_step"],
                    "config": {"workload": desc, "scheduler": "greedy", "device": info["name"]},
                    "h2d_bytes_per_step": m["h2d_bytes_per_step"], "roofline": roofline(m, "c4", best["span"]),
                    "roofline_pcie": roofline_pcie(m, h2d_peak),
                    "stats_per_step": m["per_job"],
                    "autotune": {"seconds": tune_s, "sample": "every 16th query; the 3 fastest and the fastest of each copy mode "
                                           "again on every 4th; the fastest of those with 1, 2 and 4 walkers per "
                                           "lane on every 4th",
                                 "best": best,
                                 "trials": [{k: t[k] for k in ("stage", "partition_mb", "span", "device_schedule",
                                                               "slice_chunks", "zc_batch", "pack_partitions",
                                                               "walkers_per_thread", "split", "cache_slots", "running_ms",
                                                               "h2d_bytes")}
                                            for t in trials]}})
    else:
        ekw = dict(wl.get("eng", {}))
        ekw["span"] = args.span or ekw.get("span", wl.get("span", "partition"))
        span = ekw["span"]
        vpart, vbase = part, base
        if "blocks_mb" in wl:  # the value's engine: the graph split for block slots, the same cache bytes
            vpart = lt.Partition.build_blocks(g, wl["blocks_mb"] << 20)
            vbase = dict(base, num_device_partitions=(wl["slots"] * (wl["mb"] << 20)) // vpart.bytes,
                         zerocopy_threshold=(wl["blocks_mb"] << 20) // wl["blocks_zc_div"])
        eng = lt.Engine(g, vpart, **ekw, **vbase)
        teng = lt.Engine(g, vpart, event_timers=1, **ekw, **vbase)
        run_job(teng, wl, w)
        m = measure(eng, wl, w, args.warmup, args.steps, args.workload, timed_engine=teng)
        # at most one other engine alive while a variant is measured (hardware queues)
        teng.close()
        eng.close()
        if args.workload == "c2" and not args.no_variants:  # the reference's span: walks stop at every partition boundary
            e2 = lt.Engine(g, part, span="partition", **base)
            v = measure(e2, wl, w, 1, 2, "c2/span partition")
            out["c2_variants"] = {"span_partition": {k: v[k] for k in ("steps_per_s", "ms_per_step")}}
            e2.close()
            # the reference's figure-12 ablations of the append stage, with its span (every move
            # out of the partition is appended to a queue): each walker stored from its own lane,
            # and one queue reservation per walker
            for mode, tag in ((1, "span_partition_no_random_access_reduce"), (2, "span_partition_one_level_queue")):
                e2 = lt.Engine(g, part, span="partition", append_mode=mode, **base)
                v = measure(e2, wl, w, 1, 3, f"c2/{tag}")
                out["c2_variants"][tag] = {k: v[k] for k in ("steps_per_s", "ms_per_step")}
                e2.close()
        if args.workload == "c3" and not args.no_variants:
            keep = ("steps_per_s", "ms_per_step", "median_ms_per_step", "e2e_ms_per_step", "h2d_bytes_per_step",
                    "h2d_GB_per_job", "pcie_GBps")
            variants = {}
            # the reference's rules: span partition, raw copies, default zero-copy
            # threshold; then its figure-14 ablation (explicit copies only, zero-copy only)
            for name, kw, k in (("reference_semantics", {}, args.steps), ("explicit_only", dict(zerocopy_threshold=0), 1),
                                ("zerocopy_only", dict(zerocopy_threshold=1 << 62), 1),
                                ("span_resident_zc_batch16", dict(span="resident", zc_batch=16), 2),
                                ("packed", dict(pack_partitions=1), 2),
                                ("blocks", dict(pack_partitions=2), 2),
                                ("span_resident_zc_batch16_packed", dict(span="resident", zc_batch=16,
                                                                         pack_partitions=1), 2)):
                e2 = lt.Engine(g, part, **dict(base, **kw))
                v = measure(e2, wl, w, 1, k, f"c3/{name}")
                variants[name] = {x: v[x] for x in keep}
                e2.close()
            out["c3_variants"] = variants
        out.update({"value": m["steps_per_s"], "ms_per_step": m["ms_per_step"],
                    "median_ms_per_step": m["median_ms_per_step"], "e2e_ms_per_step": m["e2e_ms_per_step"],
                    "config": {"workload": desc,
                               "scheduler": ("plain-rule prefetch (depth 2, plain victims, no network)"
                                             if ekw.get("prefetch_plain") == 1 else "greedy"),
                               **{k: v for k, v in ekw.items()}, "device": info["name"],
                               **({"partitions": f"{vpart.count} partitions sized for {wl['blocks_mb']} MiB "
                                                 f"block slots (lt_partition_build_blocks), zero-copy threshold "
                                                 f"slot/{wl['blocks_zc_div']}"} if vpart is not part else {})},
                    "h2d_bytes_per_step": m["h2d_bytes_per_step"], "roofline": roofline(m, args.workload, span),
                    **({"roofline_pcie": roofline_pcie(m, h2d_peak)} if wl["slots"] else {}),
                    "stats_per_step": m["per_job"]})
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(g, wl, w)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
