"""lt_scorer_infer_noalloc, the inference entry point that allocates nothing
(its buffers are a static per-GPU scratch), against lt_scorer_infer through
the C ABI: the same kernel on the same state and weights, so the scores must
be equal bit for bit.  Two calls in a row, the second with a smaller scorer
(fewer partitions), then back to the first, then after a weight update."""
import numpy as np
import pytest

import lighttraffic_amd as lt

pytestmark = pytest.mark.gpu


def _scorer(P, seed):
    rng = np.random.default_rng(seed)
    node = np.concatenate([[0], np.cumsum(rng.integers(1, 50, P))]).astype(np.uint32)
    edge = np.concatenate([[0], np.cumsum(rng.integers(0, 400, P))]).astype(np.uint64)
    h = lt._vp()
    lt._check(lt.lib().lt_partition_from_offsets(P, 1 << 20, node.ctypes.data, edge.ctypes.data,
                                                 lt.ctypes.byref(h)), "lt_partition_from_offsets")
    part = lt.Partition(h)
    sc = lt.Scorer(part, seed=seed)
    sc.set_params(rng.normal(0, 0.3, lt.Scorer.NUM_PARAMS).astype(np.float32))
    return sc, part


def _state(P, seed):
    rng = np.random.default_rng(seed)
    counts = (rng.integers(0, 5, P) * rng.integers(0, 100000, P)).astype(np.uint32)
    return (counts, (rng.random(P) < 0.5).astype(np.uint8), rng.integers(0, 3 * P, P).astype(np.uint32),
            rng.integers(-1, P, 8).astype(np.int32),
            int(rng.integers(0, P)), float(counts.sum() * 3 + 1))


def test_noalloc_inference_matches_the_allocating_entry_point():
    big, pbig = _scorer(429, 1)
    small, psmall = _scorer(37, 2)
    for sc, P, seed in ((big, 429, 10), (small, 37, 11), (big, 429, 12)):
        st = _state(P, seed)
        ref = sc.infer(*st)
        got = sc.infer_noalloc(*st)
        assert got.shape == (P, 2) and np.isfinite(got).all()
        assert np.array_equal(got, ref), (P, seed)
    # new weights reach the scratch's copy of them
    big.set_params(np.random.default_rng(5).normal(0, 0.3, lt.Scorer.NUM_PARAMS).astype(np.float32))
    st = _state(429, 13)
    assert np.array_equal(big.infer_noalloc(*st), big.infer(*st))
    out = np.zeros((37, 2), np.float32)  # caller-owned output
    st = _state(37, 14)
    assert small.infer_noalloc(*st, out=out) is out and np.array_equal(out, small.infer(*st))
