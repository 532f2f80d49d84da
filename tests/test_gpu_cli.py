"""The pagerank / ppr / genericwalk programs on the GPU (app/*.cu main,
src/Config.h:38-89): reference flags, the stdout lines the reference's
benchmark scripts grep (benchmark/figure13/iter_analysis.sh,
benchmark/figure9/run.sh), the result check of app/pagerank.cu:90-101 and
app/ppr.cu:78-89, and the learned scheduler driven by a saved scorer."""
import os
import re
import subprocess

import numpy as np
import pytest

import lighttraffic_amd as lt

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def graph_file(tmp_path_factory):
    d = tmp_path_factory.mktemp("cli")
    path = str(d / "s16.bcsr")
    lt.Graph.rmat(16, 16).save(path)
    return path


def run(app, *args):
    cmd = [os.path.join(lt.BIN_DIR, app), *map(str, args)]
    log(" ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


def grep_num(pattern, out):
    m = re.search(pattern, out)
    assert m, (pattern, out)
    return int(m.group(1))


def test_pagerank_cli_capped_cache(graph_file):
    out = run("pagerank", "-f", graph_file, "-m", "-s", 1, "-p", 3, "-w", 100000, "-l", 10, "--timers")
    assert "result correct" in out
    nparts = grep_num(r"number of partitons: (\d+)", out)
    assert nparts > 3
    # the three numbers benchmark/figure13/iter_analysis.sh extracts
    it, ex = grep_num(r"iterations: (\d+)", out), grep_num(r"explicit: (\d+)", out)
    loads = grep_num(r"graph loading time: [0-9.]+ ms, calls: (\d+)", out)
    assert it >= ex >= 1 and loads >= nparts - 3
    assert re.search(r"Running Time: [0-9.]+ ms", out)
    # figure 12-14 ablations as runtime flags: FIFO, no pipelining, zero copy off / always,
    # one queue reservation per walker, walkers stored from their own lanes
    for flags in (["--scheduler", "fifo"], ["--no-pipeline"], ["--zerocopy-threshold", 1],
                  ["--zerocopy-threshold", 1 << 40], ["--one-level-queue"], ["--no-random-access-reduce"]):
        out = run("pagerank", "-f", graph_file, "-p", 3, "-w", 50000, *flags)
        assert "result correct" in out
    out = run("pagerank", "-f", graph_file, "-p", 3, "-w", 50000, "--pack", "--zerocopy-threshold", 0)
    assert "result correct" in out
    assert grep_num(r"H2D bytes: explicit (\d+)", out) < grep_num(r"H2D bytes unpacked: (\d+)", out)
    out = run("pagerank", "-f", graph_file, "-p", 3, "-w", 50000, "--pack-blocks", "--zerocopy-threshold", 0)
    assert "result correct" in out
    assert grep_num(r"H2D bytes: explicit (\d+)", out) < grep_num(r"H2D bytes unpacked: (\d+)", out)
    out = run("pagerank", "-f", graph_file, "-p", 3, "-w", 50000, "--zerocopy-threshold", 1 << 40)
    assert grep_num(r"graph loading time: [0-9.]+ ms, calls: (\d+)", out) == 0
    # -b below -w: walker queues spill to host memory and come back (the
    # reference's walker loading / evicting lines, src/Scheduler.h:440-453)
    out = run("pagerank", "-f", graph_file, "-p", 3, "-w", 100000, "-b", 20000, "--timers")
    assert "result correct" in out
    assert grep_num(r"walker loading time: [0-9.]+ ms, calls: (\d+)", out) > 0
    assert grep_num(r"walker evicting time: [0-9.]+ ms, calls: (\d+)", out) > 0


def test_ppr_and_genericwalk_cli(graph_file):
    run("pagerank", "-f", graph_file, "-m", "-s", 1)  # materialise the .part file
    out = run("ppr", "-f", graph_file, "-p", 4, "-w", 200000, "--source", 7, "--prob", 0.15, "--span", "resident",
              "--zc-batch", 8, "--zerocopy-threshold", 4000)
    assert "result correct" in out
    out = run("genericwalk", "-f", graph_file, "-p", 2, "-w", 100000, "-l", 20, "--device-schedule")
    assert re.search(r"Running Time: [0-9.]+ ms", out)


def test_ppr_queries_autotuned(graph_file, tmp_path):
    """ppr --queries with --tune: lt_tune_ppr_batch picks the partition size
    and engine configuration, and every walker's endpoint is counted."""
    g = lt.Graph.open(graph_file)
    src, cnt = g.ppr_queries(64, lo=100, hi=5000, seed=3)
    qfile = tmp_path / "queries.txt"
    qfile.write_text("".join(f"{s} {c}\n" for s, c in zip(src, cnt)))
    run("pagerank", "-f", graph_file, "-m", "-s", 1)
    out = run("ppr", "-f", graph_file, "-p", 4, "--queries", qfile, "--tune")
    m = re.search(r"autotuned: (\d+) MB (?:partitions|block slots \(split for them\)), (\d+) cache slots, "
                  r"span (\w+), (\w+) scheduling", out)
    assert m, out
    assert int(m.group(1)) in (16, 32, 64) and m.group(3) in ("partition", "resident")
    assert "result correct" in out
    assert grep_num(r"sum: \d+, (\d+)", out) == int(cnt.sum())


def test_learned_scheduler_from_saved_scorer(graph_file, tmp_path):
    g = lt.Graph.open(graph_file)
    part = lt.Partition.load(graph_file)
    eng = lt.Engine(g, part, num_device_partitions=3, span="resident", record_trace=1)
    eng.pagerank(60000, 10)
    ref_moves = eng.stats()["moves"]
    sc = lt.Scorer(part, seed=3)
    sc.add_trace(eng, True, True)
    sc.fit(epochs=3)
    weights = str(tmp_path / "scorer.bin")
    sc.save(weights)
    out = run("pagerank", "-f", graph_file, "-p", 3, "-w", 60000, "-l", 10, "--scheduler", "learned",
              "--scorer", weights, "--span", "resident", "--json")
    assert "result correct" in out
    # same walkers and seed as the engine run above, so the same moves
    assert grep_num(r'"moves": (\d+)', out) == ref_moves
    # the fastest capped-cache setup: the same with packed partition copies
    out = run("pagerank", "-f", graph_file, "-p", 3, "-w", 60000, "-l", 10, "--scheduler", "learned",
              "--scorer", weights, "--span", "resident", "--pack", "--json")
    assert "result correct" in out and "H2D bytes unpacked" in out
    assert grep_num(r'"moves": (\d+)', out) == ref_moves
    # and with partitions kept packed in HBM (the bench's headline setup)
    out = run("pagerank", "-f", graph_file, "-p", 3, "-w", 60000, "-l", 10, "--scheduler", "learned",
              "--scorer", weights, "--span", "resident", "--pack-blocks", "--json")
    assert "result correct" in out and "H2D bytes unpacked" in out
    assert grep_num(r'"moves": (\d+)', out) == ref_moves
    # the learned loop's settings as flags: largest-queue prefetch targets at
    # depth 3 with the network's victims, resident batches, 1 walker per lane,
    # the network's next partition among near-ties; then the same loop with
    # plain victims and no scorer
    for extra in (["--scorer", weights, "--prefetch-plain", 2, "--learned-next", 50, "--evict-shortlist", 2],
                  ["--prefetch-plain", 1]):
        out = run("pagerank", "-f", graph_file, "-p", 3, "-w", 60000, "-l", 10, "--scheduler", "learned",
                  "--span", "resident", "--pack-blocks", "--prefetch-depth", 3, "--resident-batch", 16,
                  "--walkers-per-thread", 1, "--slice-chunks", 64, "--json", *extra)
        assert "result correct" in out
        assert grep_num(r'"moves": (\d+)', out) == ref_moves


def test_split_for_block_slots_from_the_cli(graph_file, tmp_path):
    """-m --split-blocks writes the .part of lt_partition_build_blocks;
    --cache-mb sizes the cache in bytes of that table's partition size; the
    walk moves as on the reference's split."""
    gf = str(tmp_path / "g.bcsr")
    with open(graph_file, "rb") as src, open(gf, "wb") as dst:
        dst.write(src.read())
    out = run("pagerank", "-f", gf, "-m", "-s", 1, "--split-blocks", "-w", 60000, "--cache-mb", 3,
              "--pack-blocks", "--span", "resident", "--json")
    assert "result correct" in out
    g = lt.Graph.open(gf)
    part, ref = lt.Partition.load(gf), lt.Partition.build_blocks(g, 1 << 20)
    assert np.array_equal(part.node_offsets(), ref.node_offsets()) and part.bytes == ref.bytes
    base = run("pagerank", "-f", graph_file, "-w", 60000, "--span", "resident", "--json")
    assert grep_num(r'"moves": (\d+)', out) == grep_num(r'"moves": (\d+)', base)
