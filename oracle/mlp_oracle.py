"""CPU oracle of the cut-pair surrogate network (test infrastructure only).

Restates, in numpy, the arithmetic of the BF16 MFMA kernels of libnlfem_amd.so: every
operand that the GPU feeds to a matrix instruction is rounded to BF16 (round to nearest
even), products are exact and accumulated in float64 here (FP32 on the GPU, so the two
agree to accumulation-order rounding, not bit for bit), SiLU and its derivative in float64.
With bf16=False the same code is plain float64 and is pinned by finite differences and by
torch autograd in tests/test_mlp_oracle.py.

The network has no counterpart in the reference: it replaces the inner quadrature of the
partially cut (pair, outer point) items of integrate_retriangulate
(src/integration.cpp:345-540) and is judged against that exact method.

Parameter layout (row-major W[out][in]):
    W1[64][12] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[9][64] b4[9]
norm: input mean[12], input 1/std[12], output mean[9], output std[9].
"""
import numpy as np

NIN, NH, NOUT = 12, 64, 9
SHAPES = [(NH, NIN), (NH,), (NH, NH), (NH,), (NH, NH), (NH,), (NOUT, NH), (NOUT,)]
NPARAM = sum(int(np.prod(s)) for s in SHAPES)


def bf16(x):
    """Round float values to BF16 (nearest even) and return them as float64."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params):
    out, o = [], 0
    p = np.asarray(params, dtype=np.float64)
    for s in SHAPES:
        n = int(np.prod(s))
        out.append(p[o:o + n].reshape(s))
        o += n
    return out


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params, norm, X, bf=True, keep=False):
    """Outputs in physical units [n, 9]; with keep=True also the intermediates (normalised)."""
    rnd = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(params)
    norm = np.asarray(norm, dtype=np.float64)
    mu, isd, omu, osd = norm[0:12], norm[12:24], norm[24:33], norm[33:42]
    x = (np.asarray(X, dtype=np.float32).astype(np.float64) - mu) * isd
    if bf:  # the GPU normalises in float32
        x = ((np.asarray(X, dtype=np.float32) - mu.astype(np.float32)) * isd.astype(np.float32)).astype(np.float64)
    h0 = rnd(x)
    z1 = h0 @ rnd(W1).T + b1
    h1 = rnd(_silu(z1))
    z2 = h1 @ rnd(W2).T + b2
    h2 = rnd(_silu(z2))
    z3 = h2 @ rnd(W3).T + b3
    h3 = rnd(_silu(z3))
    y = h3 @ rnd(W4).T + b4
    out = y * osd + omu
    if keep:
        return out, dict(h0=h0, z1=z1, h1=h1, z2=z2, h2=h2, z3=z3, h3=h3, y=y)
    return out


def grad(params, norm, X, T, bf=True):
    """Gradient of the mean squared error (normalised output units) over all samples, and the loss."""
    rnd = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(params)
    norm = np.asarray(norm, dtype=np.float64)
    omu, osd = norm[24:33], norm[33:42]
    _, c = forward(params, norm, X, bf=bf, keep=True)
    n = X.shape[0]
    t = (np.asarray(T, dtype=np.float64) - omu) / osd
    d = c["y"] - t
    loss = float((d * d).sum() / (n * NOUT))
    dy = 2.0 * d / (n * NOUT)
    dyb = rnd(dy)
    gW4 = dyb.T @ c["h3"]
    gb4 = dy.sum(0)
    dz3 = (dyb @ rnd(W4)) * _dsilu(c["z3"])
    dz3b = rnd(dz3)
    gW3 = dz3b.T @ c["h2"]
    gb3 = dz3.sum(0)
    dz2 = (dz3b @ rnd(W3)) * _dsilu(c["z2"])
    dz2b = rnd(dz2)
    gW2 = dz2b.T @ c["h1"]
    gb2 = dz2.sum(0)
    dz1 = (dz2b @ rnd(W2)) * _dsilu(c["z1"])
    dz1b = rnd(dz1)
    gW1 = dz1b.T @ c["h0"]
    gb1 = dz1.sum(0)
    g = np.concatenate([a.ravel() for a in (gW1, gb1, gW2, gb2, gW3, gb3, gW4, gb4)])
    return g, loss
