"""BF16 MFMA surrogate kernels (libnlfem_amd.so) against the CPU oracle oracle/mlp_oracle.py.

The oracle rounds every matrix operand to BF16 as the kernels do but accumulates in float64
(FP32 on the GPU), so agreement is to accumulation-order rounding plus the occasional BF16
value that rounds the other way: tolerances below are a few BF16 ulps (2^-8 relative).
"""
import numpy as np
import pytest

import mlp_oracle as mo
from nlfem_amd import surrogate

pytestmark = pytest.mark.gpu


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, 12)).astype(np.float32) * 3 + 1
    # smooth target the network can represent (the same function for every seed)
    A = np.random.default_rng(1234).normal(size=(12, 9)) / 4
    T = (np.tanh(X.astype(np.float64) @ A) + 0.1 * np.sin(X[:, :9])).astype(np.float32)
    return X, T


@pytest.mark.parametrize("n", [1, 16, 100, 4096])
def test_forward_matches_oracle(n):
    print(f"[mlp] forward n={n}", flush=True)
    X, T = _data(n, 1)
    p = surrogate.init_params(7)
    norm = surrogate.fit_norm(X, T) if n > 1 else surrogate.fit_norm(np.vstack([X, X + 1]), np.vstack([T, T + 1]))
    Y = surrogate.forward(p, norm, X)
    R = mo.forward(p, norm, X, bf=True)
    scale = np.max(np.abs(R - norm[24:33])) + 1e-6
    assert np.max(np.abs(Y - R)) <= 3e-2 * scale
    assert np.mean(np.abs(Y - R)) <= 2e-3 * scale


def test_gradient_matches_oracle():
    print("[mlp] gradient n=1000", flush=True)
    X, T = _data(1000, 2)
    p = surrogate.init_params(8)
    norm = surrogate.fit_norm(X, T)
    g, loss = surrogate.grad(p, norm, X, T)
    gr, lr = mo.grad(p, norm, X, T, bf=True)
    assert abs(loss - lr) <= 1e-3 * lr
    off = 0
    for shape in mo.SHAPES:
        k = int(np.prod(shape))
        a, b = g[off:off + k].astype(np.float64), gr[off:off + k]
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b))
        print(f"[mlp] block {shape} rel {rel:.2e} cos {cos:.6f}", flush=True)
        assert rel < 1e-3 and cos > 0.99999
        off += k
    # and the BF16 gradient is a close approximation of the exact float64 one
    g64, _ = mo.grad(p, norm, X, T, bf=False)
    assert np.linalg.norm(g - g64) / np.linalg.norm(g64) < 5e-2


def test_masked_tail_does_not_contribute():
    X, T = _data(100, 3)
    p = surrogate.init_params(9)
    norm = surrogate.fit_norm(X, T)
    g1, l1 = surrogate.grad(p, norm, X[:70], T[:70])
    gr, lr = mo.grad(p, norm, X[:70], T[:70], bf=True)
    assert np.linalg.norm(g1 - gr) / np.linalg.norm(gr) < 2e-2


def test_training_fits_smooth_target():
    print("[mlp] training 16k samples", flush=True)
    X, T = _data(16384, 4)
    Xv, Tv = _data(2048, 5)
    p0 = surrogate.init_params(0)
    norm = surrogate.fit_norm(X, T)
    p, nm, hist = surrogate.train(X, T, epochs=40, batch=256, lr=3e-3, seed=1, params=p0, norm=norm)
    print(f"[mlp] loss first {hist[0]:.3e} last {hist[-1]:.3e}", flush=True)
    assert np.all(np.isfinite(hist))
    assert hist[-1] < 0.02 * hist[0]
    e0 = np.mean((surrogate.forward(p0, norm, Xv) - Tv) ** 2)
    e1 = np.mean((surrogate.forward(p, norm, Xv) - Tv) ** 2)
    assert e1 < 0.02 * e0
    # the oracle agrees with the GPU on the trained network too
    R = mo.forward(p, norm, Xv, bf=True)
    assert np.mean(np.abs(surrogate.forward(p, norm, Xv) - R)) < 2e-3 * np.abs(Tv).max()


def test_invalid_arguments_raise():
    from nlfem_amd._lib import NlfemError
    X, T = _data(128, 6)
    p = surrogate.init_params(0)
    norm = surrogate.fit_norm(X, T)
    with pytest.raises(NlfemError):
        surrogate.train(X, T, epochs=1, batch=100, params=p, norm=norm)  # not a multiple of 64
    bad = norm.copy()
    bad[33] = 0.0
    with pytest.raises(NlfemError):
        surrogate.forward(p, bad, X)
