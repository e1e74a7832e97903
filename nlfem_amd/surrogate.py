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


def load(path=None, kernel="constant"):
    """Load a model written by save(): params [nnet, 9737], norm [nnet, 42], meta dict.  The
    default model is the one for `kernel` (constant, or linearPrototypeMicroelasticField)."""
    if path in (None, "default"):
        path = FIELD_MODEL if kernel == FIELD_KERNEL["function"] else DEFAULT_MODEL
    with np.load(path, allow_pickle=False) as z:
        params = np.atleast_2d(z["params"].astype(np.float32))
        norm = np.atleast_2d(z["norm"].astype(np.float32))
        meta = {k: z[k].item() if z[k].ndim == 0 else z[k] for k in z.files if k not in ("params", "norm")}
    if params.shape[1] != NPARAM or norm.shape != (params.shape[0], NNORM) or not 1 <= params.shape[0] <= 4:
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


# ----------------------------------------------------------------------------- field kernel
# The microelastic field kernel (outdim 2, src/model.cpp:140-153): per partial item the six
# symmetric 2 x 2 tensors sLP_ab = sum psi_a psi_b k(y, x) w over the polygon, 18 numbers
# (xx, xy, yy per index pair 00, 01, 02, 11, 12, 22) in the item's frame (bT's centroid on +x),
# as ratios to F_ab, the same sums over the whole of bT with the kernel's trace c / |z| (so
# every ratio lies in [-1, 1]; field_scale).  Network 0 learns the first nine, network 1 the
# last nine; the library corrects their outputs to satisfy the rotation constraints exactly
# (see project_field) and hands a pair back to the exact kernel when that correction exceeds
# tol relative to the outputs.
FIELD_MODEL = os.path.join(DATA, "surrogate_field.npz")
FIELD_KERNEL = {"function": "linearPrototypeMicroelasticField", "outputdim": 2, "fractional_s": -0.5}
# Domain of s = sqrt|det bT| / delta for the field model: the batched workload's meshes
# (h in 1/32 .. 1/256, delta in 0.05 .. 0.2) span about 0.02 .. 0.63.
FIELD_SMIN, FIELD_SMAX, FIELD_TOL = 0.02, 0.7, 4e-3
_SYM = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]


def _sym6(a, b):
    return _SYM.index((min(a, b), max(a, b)))


def project_field(Y, X):
    """The library's correction of field outputs Y [n, 18] (frame units) to the nearest
    tensors with sum_b sLP_ab R u_b = 0 for a = 0, 1, 2 (u_b = X[:, 2b:2b+2], bT's vertices
    relative to x in the frame, R the rotation by 90 degrees).  Returns (corrected Y, relative
    size of the correction max|dY| / max|Y| per item)."""
    Y = np.asarray(Y, dtype=np.float64)
    X = np.asarray(X, dtype=np.float64)
    n = Y.shape[0]
    C = np.zeros((n, 6, 18))
    vx, vy = -X[:, 1::2][:, :3], X[:, 0::2][:, :3]
    for a in range(3):
        for b in range(3):
            q = _sym6(a, b)
            C[:, 2 * a, 3 * q] += vx[:, b]
            C[:, 2 * a, 3 * q + 1] += vy[:, b]
            C[:, 2 * a + 1, 3 * q + 1] += vx[:, b]
            C[:, 2 * a + 1, 3 * q + 2] += vy[:, b]
    r = np.einsum("nij,nj->ni", C, Y)
    G = np.einsum("nij,nkj->nik", C, C)
    G += 1e-12 * np.trace(G, axis1=1, axis2=2)[:, None, None] * np.eye(6)[None]
    lam = np.linalg.solve(G, r[..., None])[..., 0]
    d = np.einsum("nij,ni->nj", C, lam)
    rel = np.abs(d).max(1) / np.maximum(np.abs(Y).max(1), 1e-30)
    return Y - d, rel


def field_scale(X):
    """F [n, 6] of the library's field_scale for features X (bT's vertices relative to x in
    delta units, frame axes): sum_i dy_i psi_a psi_b(P_i) / |y_i| over the 7-point inner rule."""
    from .quadrature import rule
    Py, dy = rule("7point")
    psi = np.stack([1 - Py[:, 0] - Py[:, 1], Py[:, 0], Py[:, 1]], 1)
    V = np.asarray(X, dtype=np.float64)[:, :6].reshape(-1, 3, 2)
    Y = V[:, None, 0, :] + Py[None, :, 0, None] * (V[:, None, 1, :] - V[:, None, 0, :]) \
        + Py[None, :, 1, None] * (V[:, None, 2, :] - V[:, None, 0, :])
    w = dy[None, :] / np.linalg.norm(Y, axis=2)
    return np.stack([(w * psi[None, :, a] * psi[None, :, b]).sum(1) for a, b in _SYM], 1)


def field_training_meshes(count=24, seed=7):
    """Meshes drawn like the batched workload's: h in {1/32, 1/64, 1/128, 1/256}, jitter +-0.2h,
    delta uniform in [0.05, min(0.2, 12 h)]."""
    from .batch import c5_meshes
    return [(m, d) for m, d, _ in c5_meshes(count, seed=seed)]


def make_field_training_set(meshes, per_mesh=100_000, salt=4242, log=None):
    """Partial cut items of the field kernel with exact FP64 targets [n, 18], drawn by the
    library's exact cut kernel (16-point outer, 7-point inner rule; items do not depend on
    the outer rule)."""
    from .config import make_configuration
    from .nlfem import Assembler
    conf = make_configuration(close="retriangulate", outer="16point", inner="7point")
    Xs, Ts = [], []
    for mesh, delta in meshes:
        k = dict(FIELD_KERNEL)
        k["horizon"] = delta
        asm = Assembler(mesh, k, conf)
        try:
            asm.assemble()
            _, _, total = asm.surrogate_samples(0, 1, salt, FIELD_SMIN, FIELD_SMAX)
            stride = max(1, total // per_mesh)
            X, T, _ = asm.surrogate_samples(int(1.2 * per_mesh) + 64, stride, salt, FIELD_SMIN, FIELD_SMAX)
        finally:
            asm.close()
        if log:
            log(f"mesh nE={mesh['nE']} delta={delta:.4f}: {total} partial items, kept {len(X)} (stride {stride})")
        Xs.append(X)
        Ts.append(T)
    return np.concatenate(Xs), np.concatenate(Ts)


def build_field(out, count=24, per_mesh=100_000, epochs=40, batch=1024, lr=3e-3, seed=0, members=2, log=print):
    """Sample, train the networks of the two halves (`members` per half, different seeds and
    sample orders) on the GPU, validate on held-out meshes, save (network h * members + m is
    member m of half h)."""
    import time
    t0 = time.perf_counter()
    meshes = field_training_meshes(count + 4, seed=7 + seed)
    X, T = make_field_training_set(meshes[:count], per_mesh, log=log)
    Xv, Tv = make_field_training_set(meshes[count:], per_mesh // 4, salt=999, log=log)
    log(f"training set {len(X)} items, validation {len(Xv)}; sampling took {time.perf_counter() - t0:.1f} s")
    P, N, H = [], [], []
    for h in range(2):
        Th = T[:, 9 * h:9 * h + 9]
        nm0 = fit_norm(X, Th)
        for m in range(members):
            t1 = time.perf_counter()
            p, nm, hist = train(X, Th, epochs=epochs, batch=batch, lr=lr, seed=seed + 101 * (2 * h + m + 1), norm=nm0)
            log(f"half {h} member {m}: loss {hist[0]:.3e} -> {hist[-1]:.3e} in {time.perf_counter() - t1:.1f} s")
            P.append(p)
            N.append(nm)
            H.append(hist)
    # ensemble mean per half and the largest disagreement of the members on a ratio
    Yh, spread = [], np.zeros(len(Xv))
    for h in range(2):
        Ys = [forward(P[h * members + m], N[h * members + m], Xv) for m in range(members)]
        Yh.append(np.mean(Ys, axis=0))
        if members > 1:
            spread = np.maximum(spread, np.abs(Ys[0] - Ys[1]).max(1))
    # compare in units of c |det bT| / delta, relative to the item's full-triangle trace scale
    F = np.repeat(field_scale(Xv), 3, axis=1)
    Y = np.concatenate(Yh, axis=1) * F
    Tv = Tv * F
    Yp, corr = project_field(Y, Xv)
    rel = spread if members > 1 else corr
    scale = F.max(1)
    err_raw = np.abs(Y - Tv).max(1) / scale
    err = np.abs(Yp - Tv).max(1) / scale
    log(f"validation: item max-error / full-item scale, mean {err_raw.mean():.2e} raw, {err.mean():.2e} corrected; "
        f"correlation(indicator, error) {np.corrcoef(rel, err)[0, 1]:.2f}")
    for tol in (1e-1, 5e-2, 2e-2, 1e-2, 5e-3, 2e-3):
        acc = rel <= tol
        if acc.any():
            log(f"validation tol {tol:.0e}: accept {acc.mean():.3f} error mean {err[acc].mean():.2e} "
                f"p99 {np.quantile(err[acc], 0.99):.2e} max {err[acc].max():.2e}")
    save(out, np.stack(P), np.stack(N), smin=FIELD_SMIN, smax=FIELD_SMAX, tol=FIELD_TOL,
         kernel=FIELD_KERNEL["function"], features="v1", train_items=len(X), train_meshes=count, epochs=epochs,
         batch=batch, lr=lr, final_loss=np.stack(H)[:, -1], val_err_mean=float(err.mean()))
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
    ap.add_argument("--field", action="store_true", help="the microelastic field kernel's model instead")
    ap.add_argument("--members", type=int, default=2, help="field model: networks per half (1 or 2)")
    a = ap.parse_args()
    if a.field:
        build_field(a.out if a.out != DEFAULT_MODEL else FIELD_MODEL, a.meshes, a.per_mesh, a.epochs, a.batch, a.lr,
                    a.seed, a.members, log=lambda m: print(m, flush=True))
    else:
        build_default(a.out, a.meshes, a.per_mesh, a.epochs, a.batch, a.lr, a.seed,
                      log=lambda m: print(m, flush=True))
