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

Residual indicator.  Two networks of this shape, trained from different initialisations and
sample orders, are evaluated on every partial item; the prediction is their mean and the
indicator is their largest disagreement on any output ratio.  A pair with an item above `tol`,
or with an item outside the training domain, is integrated by the exact retriangulation
instead.  Accepted items use M := row sums of S (the basis sums to one), so every accepted
pair keeps A * 1 = 0 exactly.

Training data come from the exact method itself: nlfem_plan_surrogate_samples draws partial
items from the FP64 cut-pair kernel of assembled plans on seeded jittered meshes
(make_training_set), and train() fits them with the library's BF16 MFMA kernels.
`python -m nlfem_amd.surrogate --out FILE` rebuilds the default model this way.
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
    """Write an ensemble (params [nnet, 9737], norm [nnet, 42]) and scalar metadata."""
    p = np.atleast_2d(_f32(params))
    n = np.atleast_2d(_f32(norm))
    if p.shape[1] != NPARAM or n.shape != (p.shape[0], NNORM):
        raise ValueError("params must be [nnet, 9737] and norm [nnet, 42]")
    np.savez(path, params=p, norm=n, **{k: np.asarray(v) for k, v in meta.items()})


def load(path=None):
    """Load a model written by save(): params [nnet, 9737], norm [nnet, 42], meta dict."""
    path = DEFAULT_MODEL if path in (None, "default") else path
    with np.load(path, allow_pickle=False) as z:
        params = np.atleast_2d(z["params"].astype(np.float32))
        norm = np.atleast_2d(z["norm"].astype(np.float32))
        meta = {k: z[k].item() if z[k].ndim == 0 else z[k] for k in z.files if k not in ("params", "norm")}
    if params.shape[1] != NPARAM or norm.shape != (params.shape[0], NNORM) or not 1 <= params.shape[0] <= 2:
        raise ValueError(f"{path}: not a surrogate model")
    return params, norm, meta


# Domain of the network input size s = sqrt|det bT| / delta (training meshes cover about
# 0.03 .. 0.3) and the default acceptance threshold on the ensemble spread.
SMIN, SMAX, TOL = 0.02, 0.2, 5e-3


def training_meshes(count=20, seed=0):
    """Seeded jittered unit-square meshes (jitter +-0.3h) with delta uniform in [0.03, 0.1] and
    n in {128, 160, 192, 256} subject to delta * n <= 16 (at most 16 elements per horizon)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        delta = float(rng.uniform(0.03, 0.1))
        ns = [n for n in (128, 160, 192, 256) if delta * n <= 16.0]
        n = int(rng.choice(ns))
        out.append((n, delta, int(rng.integers(1 << 30))))
    return out


def make_training_set(meshes, per_mesh=100_000, salt=12345, log=None):
    """Partial cut items with exact FP64 targets, drawn by the library's exact cut kernel from
    assembled plans (constant kernel, retriangulate with caps, 7-point rules)."""
    from .config import make_configuration
    from .mesh import unit_square
    from .nlfem import Assembler
    Xs, Ts = [], []
    for (n, delta, mseed) in meshes:
        mesh = unit_square(n, delta, jitter=0.3, seed=mseed)
        asm = Assembler(mesh, {"function": "constant", "horizon": delta, "outputdim": 1}, make_configuration())
        try:
            asm.assemble()
            _, _, total = asm.surrogate_samples(0, 1, salt, SMIN, SMAX)
            stride = max(1, total // per_mesh)
            X, T, _ = asm.surrogate_samples(int(1.2 * per_mesh) + 64, stride, salt, SMIN, SMAX)
        finally:
            asm.close()
        if log:
            log(f"mesh n={n} delta={delta:.4f} seed={mseed}: {total} partial items, kept {len(X)} (stride {stride})")
        Xs.append(X)
        Ts.append(T)
    return np.concatenate(Xs), np.concatenate(Ts)


def build_default(out, count=20, per_mesh=100_000, epochs=30, batch=1024, lr=3e-3, seed=0, log=print):
    """Sample, train the two-network ensemble on the GPU, validate on two held-out meshes, save."""
    import time
    t0 = time.perf_counter()
    meshes = training_meshes(count + 2, seed)
    X, T = make_training_set(meshes[:count], per_mesh, log=log)
    Xv, Tv = make_training_set(meshes[count:], per_mesh // 4, salt=999, log=log)
    log(f"training set {len(X)} items, validation {len(Xv)}; sampling took {time.perf_counter() - t0:.1f} s")
    norm = fit_norm(X, T)
    P, N, H = [], [], []
    for k in range(2):
        t1 = time.perf_counter()
        p, nm, hist = train(X, T, epochs=epochs, batch=batch, lr=lr, seed=seed + 101 * (k + 1), norm=norm)
        log(f"network {k}: loss {hist[0]:.3e} -> {hist[-1]:.3e} in {time.perf_counter() - t1:.1f} s")
        P.append(p)
        N.append(nm)
        H.append(hist)
    Y = [forward(P[k], N[k], Xv) for k in range(2)]
    mean = 0.5 * (Y[0] + Y[1])
    spread = np.abs(Y[0] - Y[1]).max(1)
    err = np.abs(mean - Tv).max(1)
    for tol in (1e-2, 5e-3, 3e-3, 2e-3):
        acc = spread <= tol
        if acc.any():
            log(f"validation tol {tol:.0e}: accept {acc.mean():.3f} item max-error mean {err[acc].mean():.2e} "
                f"p99 {np.quantile(err[acc], 0.99):.2e}")
        else:
            log(f"validation tol {tol:.0e}: accept 0")
    save(out, np.stack(P), np.stack(N), smin=SMIN, smax=SMAX, tol=TOL, kernel="constant", features="v1",
         train_items=len(X), train_meshes=count, epochs=epochs, batch=batch, lr=lr,
         final_loss=np.stack(H)[:, -1], val_err_mean=float(err.mean()))
    return out


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="rebuild the cut-pair surrogate on the GPU")
    ap.add_argument("--out", default=DEFAULT_MODEL)
    ap.add_argument("--meshes", type=int, default=20)
    ap.add_argument("--per-mesh", type=int, default=100_000)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args()
    build_default(a.out, a.meshes, a.per_mesh, a.epochs, a.batch, a.lr, a.seed,
                  log=lambda m: print(m, flush=True))
