"""Pins of the surrogate network's CPU oracle (oracle/mlp_oracle.py) and its host logic.

The oracle's backward pass is checked against central finite differences and against torch
autograd (float64, an independent implementation); its BF16 rounding against torch's
float32 -> bfloat16 conversion (round to nearest even).
"""
import numpy as np
import pytest
import torch

import mlp_oracle as mo
from nlfem_amd import surrogate


def _problem(n=24, seed=0):
    rng = np.random.default_rng(seed)
    p = surrogate.init_params(seed).astype(np.float64)
    X = rng.normal(size=(n, 12)).astype(np.float32).astype(np.float64)  # inputs are FP32 on the device
    T = rng.normal(size=(n, 9))
    norm = surrogate.fit_norm(X, T).astype(np.float64)
    return p, norm, X, T


def test_layout_sizes():
    assert mo.NPARAM == surrogate.NPARAM == 9737
    p = surrogate.init_params(3)
    assert p.dtype == np.float32 and p.shape == (9737,)
    assert np.array_equal(p, surrogate.init_params(3))
    n = surrogate.fit_norm(np.ones((5, 12)), np.zeros((5, 9)))
    assert n.shape == (42,) and np.all(n[12:24] == 1) and np.all(n[33:] == 1)


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.normal(size=5000).astype(np.float32) * 10.0 ** rng.integers(-6, 6, 5000),
                        np.float32([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 0.0, 65504.0])])
    ref = torch.tensor(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(mo.bf16(x), ref)


def test_backward_finite_differences():
    p, norm, X, T = _problem()
    g, loss = mo.grad(p, norm, X, T, bf=False)

    def L(q):
        return mo.grad(q, norm, X, T, bf=False)[1]

    rng = np.random.default_rng(2)
    idx = np.concatenate([rng.choice(mo.NPARAM, 40, replace=False), [0, 767, 768, 9727, 9736]])
    for i in idx:
        e = np.zeros_like(p)
        h = 1e-5 * max(1.0, abs(p[i]))
        e[i] = h
        fd = (L(p + e) - L(p - e)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1e-3, abs(g[i])) + 1e-9, (i, fd, g[i])


def test_backward_matches_torch_autograd():
    p, norm, X, T = _problem(n=33, seed=4)
    g, loss = mo.grad(p, norm, X, T, bf=False)
    W = [torch.tensor(a, dtype=torch.float64, requires_grad=True) for a in mo.unpack(p)]
    x = (torch.tensor(X) - torch.tensor(norm[:12])) * torch.tensor(norm[12:24])
    h = x
    for li in range(3):
        h = torch.nn.functional.silu(h @ W[2 * li].T + W[2 * li + 1])
    y = h @ W[6].T + W[7]
    t = (torch.tensor(T) - torch.tensor(norm[24:33])) / torch.tensor(norm[33:42])
    lt = ((y - t) ** 2).mean()
    lt.backward()
    gt = torch.cat([w.grad.reshape(-1) for w in W]).numpy()
    assert abs(lt.item() - loss) < 1e-12 * max(1.0, loss)
    assert np.allclose(g, gt, rtol=1e-10, atol=1e-14)


def test_forward_bf16_close_to_fp64():
    p, norm, X, T = _problem(n=64, seed=5)
    y64 = mo.forward(p, norm, X, bf=False)
    y16 = mo.forward(p, norm, X, bf=True)
    # BF16 operands: relative error of a few 2^-8
    assert np.max(np.abs(y16 - y64)) <= 0.05 * np.max(np.abs(y64))
