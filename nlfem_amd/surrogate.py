"""Cut-pair surrogate: a 12 -> 64 -> 64 -> 64 -> 9 network evaluated and trained by the
library's BF16 MFMA kernels (libnlfem_amd.so, nlfem_mlp_*).

What it predicts.  In integrate_retriangulate (reference src/integration.cpp:345-540) every
outer point x_k of the outer element contributes, through the polygon B(x_k, delta) n bT, the
inner sums

    M_b  = sum_{fan triangles} sum_i w_i psi_b(y_i) k(x_k, y_i)               (3 values)
    S_ab = sum_{fan triangles} sum_i w_i psi_a(y_i) psi_b(y_i) k(x_k, y_i)    (6 values, a <= b)

which build termLocal (from sum_b M_b), termNonloc (from M) and termLocalPrime (from S).  For
the constant kernel these are polygon moments.  When bT lies entirely inside the ball they
equal the quadrature of the whole triangle, and when it does not meet the ball they are 0.
Only the partially cut items are left to the network.  It sees the item in a canonical frame:
bT's vertices relative to x_k, divided by delta, rotated so that bT's centroid lies on the
+x axis.  It returns the ratios s_ab = S_ab / S_ab(full triangle) and m_b = M_b / M_b(full).

The residual indicator is the exact identity sum_b S_ab = M_a (the basis sums to one): the
network predicts both sides, and a pair whose items violate it by more than `tol` (in units
of the full-triangle moment) is integrated by the exact retriangulation instead.  Accepted
items use M := row sums of S, so each accepted pair still has A * 1 = 0.
"""
import os

import numpy as np

from . import _lib

NIN, NH, NOUT = 12, 64, 9
NPARAM = NH * NIN + NH + 2 * (NH * NH + NH) + NOUT * NH + NOUT
NNORM = 42
DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
DEFAULT_MODEL = os.path.join(DATA, "surrogate_constant.npz")


def _f32(a, shape=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    if shape is not None and a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


def _fp(a):
    return a.ctypes.data_as(_lib._fp)


def init_params(seed=0):
    """Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights and biases, FP32."""
    rng = np.random.default_rng(seed)
    parts = []
    for fo, fi in ((NH, NIN), (NH, NH), (NH, NH), (NOUT, NH)):
        bound = 1.0 / np.sqrt(fi)
        parts.append(rng.uniform(-bound, bound, size=fo * fi))
        parts.append(rng.uniform(-bound, bound, size=fo))
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == NPARAM
    return p


def fit_norm(X, T):
    """Normalisation block: input mean and 1/std, output mean and std."""
    X = np.asarray(X, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    sd = X.std(0)
    sd[sd < 1e-12] = 1.0
    osd = T.std(0)
    osd[osd < 1e-12] = 1.0
    return np.concatenate([X.mean(0), 1.0 / sd, T.mean(0), osd]).astype(np.float32)


def forward(params, norm, X):
    """Network outputs [n, 9] computed on the GPU."""
    X = _f32(X)
    if X.ndim != 2 or X.shape[1] != NIN:
        raise ValueError("X must be [n, 12]")
    Y = np.zeros((X.shape[0], NOUT), dtype=np.float32)
    p, nm = _f32(params, (NPARAM,)), _f32(norm, (NNORM,))
    _lib.check(_lib.lib().nlfem_mlp_forward(_fp(p), _fp(nm), _fp(X), X.shape[0], _fp(Y)))
    return Y


def grad(params, norm, X, T):
    """Gradient of the mean squared normalised error over all samples (GPU), and the loss."""
    X, T = _f32(X), _f32(T)
    if X.ndim != 2 or X.shape[1] != NIN or T.shape != (X.shape[0], NOUT):
        raise ValueError("X must be [n, 12] and T [n, 9]")
    g = np.zeros(NPARAM, dtype=np.float32)
    loss = np.zeros(1, dtype=np.float32)
    p, nm = _f32(params, (NPARAM,)), _f32(norm, (NNORM,))
    _lib.check(_lib.lib().nlfem_mlp_grad(_fp(p), _fp(nm), _fp(X), _fp(T), X.shape[0], _fp(g), _fp(loss)))
    return g, float(loss[0])


def train(X, T, epochs=30, batch=1024, lr=3e-3, seed=0, params=None, norm=None):
    """Train on the GPU.  Returns (params, norm, loss per epoch)."""
    X, T = _f32(X), _f32(T)
    if X.ndim != 2 or X.shape[1] != NIN or T.shape != (X.shape[0], NOUT):
        raise ValueError("X must be [n, 12] and T [n, 9]")
    p = init_params(seed) if params is None else _f32(params, (NPARAM,)).copy()
    nm = fit_norm(X, T) if norm is None else _f32(norm, (NNORM,))
    hist = np.zeros(epochs, dtype=np.float32)
    _lib.check(_lib.lib().nlfem_mlp_train(_fp(p), _fp(nm), _fp(X), _fp(T), X.shape[0], int(epochs), int(batch),
                                          float(lr), int(seed) & 0xFFFFFFFF, _fp(hist)))
    return p, nm, hist


def save(path, params, norm, **meta):
    np.savez(path, params=_f32(params, (NPARAM,)), norm=_f32(norm, (NNORM,)),
             **{k: np.asarray(v) for k, v in meta.items()})


def load(path=None):
    """Load a trained model (params, norm, meta dict) written by save()."""
    path = DEFAULT_MODEL if path in (None, "default") else path
    with np.load(path, allow_pickle=False) as z:
        params = z["params"].astype(np.float32)
        norm = z["norm"].astype(np.float32)
        meta = {k: z[k].item() if z[k].ndim == 0 else z[k] for k in z.files if k not in ("params", "norm")}
    if params.shape != (NPARAM,) or norm.shape != (NNORM,):
        raise ValueError(f"{path}: not a surrogate model")
    return params, norm, meta
