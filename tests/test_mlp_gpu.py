"""GPU parity of the cost model's network (qfx_mlp_*) against the CPU
restatement oracle/mlp.py (float64 with the same BF16 roundings), whose own
gradient is pinned by finite differences in tests/test_oracle.py."""
import numpy as np
import pytest

from oracle import mlp
from qflex_amd import capi

pytestmark = pytest.mark.gpu


def relerr(x, ref):
    return float(np.linalg.norm(np.asarray(x) - ref) / max(np.linalg.norm(ref), 1e-30))


def case(seed, B):
    rng = np.random.default_rng(seed)
    p = mlp.init_params(seed)
    p[mlp.N_IN * mlp.N_HID:mlp.N_IN * mlp.N_HID + mlp.N_HID] = rng.standard_normal(mlp.N_HID) * 0.1
    X = rng.standard_normal((B, mlp.N_IN)).astype(np.float32)
    t = rng.standard_normal(B).astype(np.float32)
    return p, X, t


@pytest.mark.parametrize("B", [1, 37, 256, 1000])
def test_mlp_forward_and_gradients_vs_oracle(ctx, B):
    p, X, t = case(B, B)
    P, DX, DT = ctx.upload(p), ctx.upload(X), ctx.upload(t)
    Y, G, L = ctx.alloc(4 * B), ctx.alloc(4 * mlp.N_PARAMS), ctx.alloc(4)
    ctx.mlp_forward(B, DX, P, Y)
    y = ctx.download(Y, np.float32, (B,))
    y_ref, _ = mlp.forward(p, X, rounding=True)
    # FP32 accumulation vs float64; BF16-rounded activations may land on
    # the other side of a rounding boundary
    assert relerr(y, y_ref) < 2e-3, relerr(y, y_ref)
    ctx.mlp_backward(B, DX, DT, P, G, L)
    g = ctx.download(G, np.float32, (mlp.N_PARAMS,)).astype(np.float64)
    loss = float(ctx.download(L, np.float32, (1,))[0])
    loss_ref, g_ref = mlp.loss_and_grads(p, X, t, rounding=True)
    assert abs(loss - loss_ref) <= 2e-3 * loss_ref
    o = 0
    for name, shape in mlp.SIZES:
        n = int(np.prod(shape))
        err = relerr(g[o:o + n], g_ref[o:o + n])
        assert err < 2e-2, (B, name, err)
        o += n
    for b in (P, DX, DT, Y, G, L):
        b.free()


def test_mlp_adam_training_reduces_loss(ctx):
    # a smooth target of the features: the network must fit it
    rng = np.random.default_rng(0)
    B = 2048
    X = rng.standard_normal((B, mlp.N_IN)).astype(np.float32)
    t = (np.tanh(X[:, :8].sum(axis=1) / 3) + 0.1 * X[:, 8]).astype(np.float32)
    p = mlp.init_params(5)
    P, DX, DT = ctx.upload(p), ctx.upload(X), ctx.upload(t)
    G, L = ctx.alloc(4 * mlp.N_PARAMS), ctx.alloc(4)
    M = ctx.upload(np.zeros(mlp.N_PARAMS, np.float32))
    V = ctx.upload(np.zeros(mlp.N_PARAMS, np.float32))
    losses = []
    for step in range(1, 201):
        ctx.mlp_backward(B, DX, DT, P, G, L)
        if step in (1, 200):
            losses.append(float(ctx.download(L, np.float32, (1,))[0]))
        ctx.mlp_adam(P, G, M, V, step, 1e-3)
    assert losses[1] < 0.1 * losses[0], losses
    # the GPU-trained parameters, evaluated by the oracle, agree
    p_trained = ctx.download(P, np.float32, (mlp.N_PARAMS,))
    y_ref, _ = mlp.forward(p_trained, X, rounding=True)
    assert float(np.mean((y_ref - t) ** 2)) < 0.1 * losses[0]
    for b in (P, DX, DT, G, L, M, V):
        b.free()
