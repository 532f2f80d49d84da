"""FP64 restatement of the learned partition scorer (test infrastructure only).

Imported by tests/ to check the HIP training and inference kernels of
lighttraffic_amd/csrc/lt_mlp.hip; never used by the engine.  The scorer has
no counterpart in the reference; this file restates the engine's documented
model: 16 per-partition features -> 256 -> 256 -> 2 with ReLU, weights used
as their BF16 rounding (round to nearest even), all arithmetic in float64.
Gradients are taken with respect to the BF16-rounded weights (the update is
applied to FP32 masters), and are pinned by central finite differences in
the tests.
"""
import numpy as np

F, H, O, R = 16, 256, 2, 8
SIZES = [("W1", (F, H)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, O)), ("b3", (O,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(flat):
    out, off = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(flat[off:off + n], np.float64).reshape(shape)
        off += n
    return out


def pack(d):
    return np.concatenate([d[name].ravel() for name, _ in SIZES])


def bf16_round(x):
    """Round float32 values to the nearest BF16 (ties to even), returned as float64."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    nan = (u & 0x7F800000) == 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan & ((u & 0x007FFFFF) != 0), u | 0x400000, r)
    return (r.astype(np.uint32)).view(np.float32).astype(np.float64)


def shares(node_off, edge_off):
    node_off = np.asarray(node_off, np.float64)
    edge_off = np.asarray(edge_off, np.float64)
    P = len(node_off) - 1
    V = max(node_off[-1], 1.0)
    E = max(edge_off[-1], 1.0)
    vsh = np.float32(np.diff(node_off) / V * P).astype(np.float64)
    esh = np.float32(np.diff(edge_off) / E * P).astype(np.float64)
    return vsh, esh


def features(counts, resident, age, recent, cur, w_total, vsh, esh):
    """Features of one decision, shape (P, 16): queue length (three scalings),
    residency, chosen now, loaded 0..6 decisions ago, age a/(a+P) (a = walks
    of resident partitions since the partition's own last one), vertex and
    edge shares, queued walkers / W."""
    c = np.asarray(counts, np.float64)
    P = len(c)
    cmax = max(c.max(), 1.0)
    csum = c.sum()
    W = max(float(w_total), 1.0)
    mean = max(csum / P, 1.0)
    X = np.zeros((P, F))
    X[:, 0] = np.log1p(c) / np.log1p(W)
    X[:, 1] = c / cmax
    X[:, 2] = np.log1p(c / mean)
    X[:, 3] = np.asarray(resident) != 0
    X[:, 4] = np.arange(P) == cur
    for k in range(R - 1):
        X[:, 5 + k] = np.arange(P) == recent[k]
    a = np.asarray(age, np.float64)
    X[:, 12] = a / (a + P)
    X[:, 13] = vsh
    X[:, 14] = esh
    X[:, 15] = csum / W
    return X


def forward(Wb, X):
    Z1 = X @ Wb["W1"] + Wb["b1"]
    A1 = np.maximum(Z1, 0)
    Z2 = A1 @ Wb["W2"] + Wb["b2"]
    A2 = np.maximum(Z2, 0)
    Y = A2 @ Wb["W3"] + Wb["b3"]
    return A1, A2, Y


def loss_and_grad(params_flat, samples, vsh, esh, rounded=True, want_grad=True):
    """Mean losses over the batch and the gradient w.r.t. the (rounded) weights.

    samples: list of dicts with counts, resident, age, recent, cur, w_total,
    next (int, -1 = none), keep (P floats), wa, wb.
    """
    Wf = unpack(params_flat)
    Wb = {k: (bf16_round(v) if rounded else v) for k, v in Wf.items()}
    B = len(samples)
    grads = {k: np.zeros_like(v) for k, v in Wb.items()}
    la = lb = 0.0
    for s in samples:
        X = features(s["counts"], s["resident"], s["age"], s["recent"], s["cur"], s["w_total"], vsh, esh)
        P = X.shape[0]
        A1, A2, Y = forward(Wb, X)
        dY = np.zeros_like(Y)
        lab = s["next"]
        c = np.asarray(s["counts"])
        wa = s["wa"] if lab >= 0 else 0.0
        cand = (c > 0) | (np.arange(P) == lab)
        if wa > 0 and cand.any():
            z = Y[cand, 0]
            m = z.max()
            e = np.exp(z - m)
            se = e.sum()
            p = np.zeros(P)
            p[cand] = e / se
            onehot = np.zeros(P)
            onehot[lab] = 1.0
            la += wa * (np.log(se) + m - Y[lab, 0]) / B
            dY[:, 0] = np.where(cand, wa * (p - onehot) / B, 0.0)
        d = Y[:, 1] - np.asarray(s["keep"], np.float64)
        lb += s["wb"] * np.sum(d * d) / B / P
        dY[:, 1] = s["wb"] * 2 * d / B / P
        if not want_grad:
            continue
        grads["W3"] += A2.T @ dY
        grads["b3"] += dY.sum(0)
        dA2 = (dY @ Wb["W3"].T) * (A2 > 0)
        grads["W2"] += A1.T @ dA2
        grads["b2"] += dA2.sum(0)
        dA1 = (dA2 @ Wb["W2"].T) * (A1 > 0)
        grads["W1"] += X.T @ dA1
        grads["b1"] += dA1.sum(0)
    return np.array([la, lb]), pack(grads)
