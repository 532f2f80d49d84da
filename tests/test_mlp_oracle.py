"""Pins of the FP64 scorer oracle (oracle/mlp_ref.py).

- BF16 rounding against torch's own float32 -> bfloat16 conversion (round to
  nearest even) and hand-checked ties;
- analytic gradients against central finite differences of the loss in
  float64, with respect to the BF16-rounded weights;
- features against values computed by hand for a small decision.
"""
import numpy as np
import pytest
import torch

from oracle import mlp_ref as M


def test_bf16_round_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(10000).astype(np.float32) * 10.0 ** rng.integers(-8, 8, 10000),
                        np.array([0.0, -0.0, 1.0, 1.00390625, 1.01171875, 65504.0, 3e38], np.float32)])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(M.bf16_round(x), ref)
    assert M.bf16_round(np.float32(1.00390625)) == 1.0          # tie, even stays
    assert M.bf16_round(np.float32(1.01171875)) == 1.015625     # tie, odd rounds up


def _sample(rng, P, with_next=True):
    counts = (rng.integers(0, 5, P) * rng.integers(0, 10000, P)).astype(np.uint32)
    lab = int(rng.integers(0, P)) if with_next else -1
    return {"counts": counts, "resident": (rng.random(P) < 0.5).astype(np.uint8),
            "age": rng.integers(0, 3 * P, P).astype(np.uint32),
            "recent": rng.integers(-1, P, 8).astype(np.int32), "cur": int(rng.integers(0, P)),
            "w_total": float(counts.sum() * 2 + 1), "next": lab, "keep": rng.random(P).astype(np.float32) * 2,
            "wa": 1.0, "wb": 1.0}


def test_features_by_hand():
    counts = np.array([0, 10, 30, 0], np.uint32)
    X = M.features(counts, [1, 0, 0, 1], [0, 4, 12, 1], [2, -1, -1, -1, -1, -1, 3, 0], 1, 100.0,
                   np.array([1.0, 2.0, 0.5, 0.5]), np.array([0.25, 1.0, 2.0, 0.75]))
    assert X.shape == (4, 16)
    assert np.allclose(X[:, 0], np.log1p([0, 10, 30, 0]) / np.log1p(100))
    assert np.allclose(X[:, 1], [0, 1 / 3, 1, 0])
    assert np.allclose(X[:, 2], np.log1p(np.array([0, 10, 30, 0]) / 10.0))
    assert list(X[:, 3]) == [1, 0, 0, 1] and list(X[:, 4]) == [0, 1, 0, 0]
    assert list(X[:, 5]) == [0, 0, 1, 0] and list(X[:, 11]) == [0, 0, 0, 1]
    assert np.all(X[:, 6:11] == 0)  # the 8th recent load (partition 0) is not a feature
    assert list(X[:, 12]) == [0, 0.5, 0.75, 0.2]  # a / (a + P), P = 4
    assert np.allclose(X[:, 15], 0.4)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    P = 7
    vsh, esh = M.shares(np.array([0, 3, 5, 9, 10, 14, 20, 21]), np.array([0, 30, 31, 60, 80, 90, 99, 120]))
    params = rng.standard_normal(M.NPARAMS) * 0.2
    params = M.bf16_round(params.astype(np.float32))  # exact BF16 values: rounding is the identity
    samples = [_sample(rng, P), _sample(rng, P, with_next=False), _sample(rng, P)]
    loss, g = M.loss_and_grad(params, samples, vsh, esh, rounded=False)
    idx = np.concatenate([rng.integers(0, M.NPARAMS, 60), np.arange(M.NPARAMS - 600, M.NPARAMS, 37)])
    h = 1e-6
    for i in idx:
        p1 = params.copy(); p1[i] += h
        p2 = params.copy(); p2[i] -= h
        l1, _ = M.loss_and_grad(p1, samples, vsh, esh, rounded=False, want_grad=False)
        l2, _ = M.loss_and_grad(p2, samples, vsh, esh, rounded=False, want_grad=False)
        fd = (l1.sum() - l2.sum()) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, g[i])
