"""CPU restatement of the cost model's network (test infrastructure only).

The GPU network (qflex_amd/csrc/qfx_mlp.hip) is an MLP 64 -> 256 -> 256 -> 1
with ReLU whose GEMM operands are rounded to BF16 (round to nearest even)
and accumulated in FP32, trained on the mean squared error.  This module
computes the same forward and backward passes in float64, with the BF16
roundings at the same places (``rounding=True``) or without them
(``rounding=False``: the plain float64 network, whose gradient is pinned
by central finite differences in tests/test_oracle.py).

There is no reference counterpart: the network scores candidate
contraction orderings.
"""
import numpy as np

N_IN, N_HID = 64, 256
SIZES = [("W1", (N_IN, N_HID)), ("b1", (N_HID,)), ("W2", (N_HID, N_HID)), ("b2", (N_HID,)),
         ("w3", (N_HID,)), ("b3", (1,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(params):
    """Views of the flat FP32 parameter vector, in the C-ABI's layout."""
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(params[o:o + n]).reshape(shape)
        o += n
    return out


def bf16(x):
    """float32 -> nearest BF16 (ties to even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def init_params(seed):
    """He-style initial weights, zero biases (FP32)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS, dtype=np.float32)
    v = unpack(p)
    v["W1"][:] = rng.standard_normal((N_IN, N_HID)) * np.sqrt(2.0 / N_IN)
    v["W2"][:] = rng.standard_normal((N_HID, N_HID)) * np.sqrt(2.0 / N_HID)
    v["w3"][:] = rng.standard_normal(N_HID) * np.sqrt(1.0 / N_HID)
    return p


def forward(params, X, rounding=True):
    r = bf16 if rounding else (lambda t: np.asarray(t, dtype=np.float64))
    P = unpack(np.asarray(params, dtype=np.float32))
    Xb = r(X)
    H1 = np.maximum(Xb @ r(P["W1"]) + P["b1"].astype(np.float64), 0.0)
    H1b = r(H1.astype(np.float32)) if rounding else H1
    H2 = np.maximum(H1b @ r(P["W2"]) + P["b2"].astype(np.float64), 0.0)
    H2b = r(H2.astype(np.float32)) if rounding else H2
    y = H2b @ r(P["w3"]) + float(P["b3"][0])
    return y, (Xb, H1b, H2b)


def loss_and_grads(params, X, t, rounding=True):
    """Mean squared error and its gradient (flat, C-ABI layout)."""
    r = bf16 if rounding else (lambda v: np.asarray(v, dtype=np.float64))
    P = unpack(np.asarray(params, dtype=np.float32))
    y, (Xb, H1b, H2b) = forward(params, X, rounding)
    B = len(t)
    d = y - np.asarray(t, dtype=np.float64)
    loss = float(np.mean(d * d))
    dy = 2.0 * d / B
    g = {}
    g["w3"] = H2b.T @ (r(dy.astype(np.float32)) if rounding else dy)
    g["b3"] = np.array([dy.sum()])
    dZ2 = (H2b > 0) * dy[:, None] * r(P["w3"])[None, :]
    dZ2b = r(dZ2.astype(np.float32)) if rounding else dZ2
    g["W2"] = H1b.T @ dZ2b
    g["b2"] = dZ2.sum(axis=0)
    dH1 = dZ2b @ r(P["W2"]).T
    dZ1 = (H1b > 0) * dH1
    dZ1b = r(dZ1.astype(np.float32)) if rounding else dZ1
    g["W1"] = Xb.T @ dZ1b
    g["b1"] = dZ1.sum(axis=0)
    flat = np.concatenate([np.asarray(g[name], dtype=np.float64).reshape(-1) for name, _ in SIZES])
    return loss, flat


def finite_difference(params, X, t, idx, h=1e-4):
    """Central difference of the float64 loss along parameter idx."""
    p = np.asarray(params, dtype=np.float64).copy()

    def loss_at(v):
        q = p.copy()
        q[idx] = v
        y, _ = forward_f64(q, X)
        d = y - np.asarray(t, dtype=np.float64)
        return float(np.mean(d * d))

    return (loss_at(p[idx] + h) - loss_at(p[idx] - h)) / (2 * h)


def forward_f64(params, X):
    """The network in float64 throughout (no FP32 storage of parameters)."""
    o, P = 0, {}
    for name, shape in SIZES:
        n = int(np.prod(shape))
        P[name] = np.asarray(params[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    X = np.asarray(X, dtype=np.float64)
    H1 = np.maximum(X @ P["W1"] + P["b1"], 0.0)
    H2 = np.maximum(H1 @ P["W2"] + P["b2"], 0.0)
    return H2 @ P["w3"] + P["b3"][0], (X, H1, H2)
