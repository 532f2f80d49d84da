"""Benchmark of the MI355X random-walk engine (one JSON line on stdout).

A step is one complete walk job over one synthetic R-MAT graph: every walker
created, walked to its end under the partition scheduler, results reduced.
Timed region per step = the scheduler loop ("Running Time" of
src/Scheduler.h:425-441): graph generation, walker creation and the copy of
results back to the host are outside it, as in the reference.

Workloads (R-MAT a,b,c,d = .57/.19/.19/.05, seeded, deduped, symmetrised,
self loops and isolated vertices removed):
  c1  scale 16 ef 16, one 32 MB partition, pagerank w=2|V| l=10
  c2  scale 22 ef 16, 32 MB partitions all HBM-resident, genericwalk w=2|V| l=10
  c3  scale 25 ef 44, 4 GB partition cache, ppr from vertex 0 w=2|V| p=.15
  c5  scale 26 ef 27, 8 GB partition cache, pagerank w=2|V| l=10
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "walk steps/s & H2D bytes/step: learned vs plain scheduler, 1 GPU"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "c1": dict(scale=16, ef=16, mb=32, app="pagerank", slots=0, l=10),
    "c2": dict(scale=22, ef=16, mb=32, app="genericwalk", slots=0, l=10),
    "c3": dict(scale=25, ef=44, mb=32, app="ppr", slots=128, l=0, wdev=1 << 26),
    "c5": dict(scale=26, ef=27, mb=32, app="pagerank", slots=256, l=10),
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes(st):
    """rowptr pair 16 B + col 4 B per move, a u32 read-modify-write per visit,
    one 16-byte walker record read and written per queue residency."""
    return 20 * st["moves"] + 8 * st["visits"] + 32 * st["residencies"]


def run_job(eng, wl, w):
    if wl["app"] == "pagerank":
        eng.pagerank(w, wl["l"])
    elif wl["app"] == "genericwalk":
        eng.genericwalk(w, wl["l"], endpoints=True)
    else:
        eng.ppr(w, source=0, prob=0.15)
    return eng.stats()


def cpu_baseline(g, wl, w, budget_s=20.0):
    """The oracle (OpenMP, whole graph in DRAM) on a bounded sample of the
    same workload; steps/s scaled to the sample."""
    from oracle import ffi as oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    rowptr, col = g.rowptr(), g.col()
    n = min(w, 200000)
    while True:
        t0 = time.perf_counter()
        if wl["app"] == "pagerank":
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
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS))
    ap.add_argument("--span", default="partition", choices=["partition", "resident"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU sharding is a design note only",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import lighttraffic_amd as lt

    wl = WORKLOADS[args.workload]
    info = lt.device_info()
    t0 = time.perf_counter()
    g = lt.Graph.rmat(wl["scale"], wl["ef"])
    log(f"graph s{wl['scale']} ef{wl['ef']}: V={g.num_node} E={g.num_edge} in {time.perf_counter() - t0:.1f} s")
    part = lt.Partition.build(g, wl["mb"] << 20)
    w = 2 * g.num_node
    cfg = dict(num_device_partitions=wl["slots"], event_timers=1, span=args.span)
    if "wdev" in wl:
        cfg["num_device_walkers"] = wl["wdev"]
    eng = lt.Engine(g, part, **cfg)
    for i in range(args.warmup):
        st = run_job(eng, wl, w)
        log(f"warmup {i}: {st['running_ms']:.2f} ms")
    stats = []
    for i in range(args.steps):
        stats.append(run_job(eng, wl, w))
        log(f"step {i}: {stats[-1]['running_ms']:.2f} ms, launches {stats[-1]['kernel_launches']}")
    ms = float(np.mean([s["running_ms"] for s in stats]))
    nominal = stats[0]["nominal_steps"]
    tot = {k: sum(s[k] for s in stats) for k in ("moves", "visits", "residencies", "kernel_launches",
                                                  "h2d_bytes_explicit", "h2d_bytes_zerocopy")}
    kernel_ms = sum(s["walk_kernel_ms"] for s in stats)
    launches = tot["kernel_launches"]
    alg_bytes = algorithmic_bytes(tot)
    achieved = alg_bytes / (kernel_ms * 1e-3) / 1e9 if kernel_ms > 0 else 0.0
    out = {
        "metric": METRIC,
        "value": nominal / (ms * 1e-3),
        "unit": "walk steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic R-MAT (a,b,c,d=.57/.19/.19/.05), generated on the GPU",
        "config": {"workload": f"{args.workload}: R-MAT scale {wl['scale']} ef {wl['ef']}, {wl['app']} "
                               f"w=2|V|={w}" + (f" l={wl['l']}" if wl['l'] else " p=0.15") +
                               f", {wl['mb']} MB partitions, cache "
                               + ("all" if not wl["slots"] else f"{wl['slots']} slots"),
                   "V": g.num_node, "E": g.num_edge, "partitions": part.count, "scheduler": "greedy",
                   "span": args.span, "device": info["name"]},
        "h2d_bytes_per_step": (tot["h2d_bytes_explicit"] + tot["h2d_bytes_zerocopy"]) / (nominal * len(stats)),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "walk_kernel", "launches": launches,
                     "avg_launch_us": kernel_ms * 1e3 / max(launches, 1),
                     "algorithmic_bytes_per_launch": alg_bytes / max(launches, 1)},
        "stats_per_step": {k: v / len(stats) for k, v in tot.items()},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(g, wl, w)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
