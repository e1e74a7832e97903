"""Cut-pair surrogate inside the assembly (k_cut_sur + exact fallback), against the FP64 oracle.

Bars: sparsity pattern identical; A * 1 = 0 to rounding (accepted pairs keep the partition
of unity); matrix relative Frobenius error of the default model below 2e-5 on an unseen
mesh; with tol = 0 every pair with a partially cut item goes back to the exact method and the
result matches the oracle to the FP64 bar (1e-12 of the sum of |addends|).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from nlfem_amd import nlfem, surrogate
from nlfem_amd.config import make_configuration
from nlfem_amd.mesh import unit_square
from nlfem_amd.quadrature import rule

pytestmark = pytest.mark.gpu


def _assemble(mesh, kernel, conf, sur=None):
    a = nlfem.Assembler(mesh, kernel, conf)
    try:
        if sur is not None:
            a.set_surrogate(**sur)
        a.assemble()
        return a.csr(), a.stats()
    finally:
        a.close()


def _case(n=48, delta=0.2, jitter=0.2, seed=11):
    mesh = unit_square(n, delta, jitter=jitter, seed=seed)
    return mesh, {"function": "constant", "horizon": delta, "outputdim": 1}, make_configuration()


def test_samples_are_exact_moments():
    """Targets drawn by the GPU equal the oracle's inner sums for the geometry the features encode."""
    mesh, kernel, conf = _case()
    a = nlfem.Assembler(mesh, kernel, conf)
    try:
        a.assemble()
        F, T, n = a.surrogate_samples(400, 97, 5, surrogate.SMIN, surrogate.SMAX)
    finally:
        a.close()
    print(f"[sur] samples kept {len(F)} of {n}", flush=True)
    assert len(F) > 100
    Px, dx = np.array([[1 / 3, 1 / 3]]), np.array([1.0])
    Py, dy = rule("7point")
    psi = np.stack([1 - Py[:, 0] - Py[:, 1], Py[:, 0], Py[:, 1]], 1)
    q2 = np.array([(dy * psi[:, i] * psi[:, j]).sum() for i, j in ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))])
    q1 = (dy[:, None] * psi).sum(0)
    eps = 1e-4
    Aout = np.array([[-eps, -eps], [2 * eps, -eps], [-eps, 2 * eps]])  # centroid at the origin
    worst = 0.0
    for f, t in zip(F[:120], T[:120]):
        Bv = f[:6].astype(np.float64).reshape(3, 2)  # canonical frame, delta = 1, outer point at 0
        verts = np.concatenate([Aout, Bv])
        r, bl = oracle.integrate_pair("retriangulate", np.array([[0, 1, 2], [3, 4, 5]]), verts, 0, 1, Px, dx,
                                      Py, dy, 1.0, "constant")
        detA = abs((Aout[1, 0] - Aout[0, 0]) * (Aout[2, 1] - Aout[0, 1]) - (Aout[2, 0] - Aout[0, 0]) * (Aout[1, 1] - Aout[0, 1]))
        detB = abs((Bv[1, 0] - Bv[0, 0]) * (Bv[2, 1] - Bv[0, 1]) - (Bv[2, 0] - Bv[0, 0]) * (Bv[1, 1] - Bv[0, 1]))
        c = 4 / np.pi
        S = bl[2] / (detA * c * detB)  # termLocalPrime of one outer point with weight 1
        s = np.array([S[i, j] for i, j in ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))]) / q2
        m = bl[1].sum(0) / (detA * c * detB) / q1 / 1.0  # column sums of termNonloc: sum_a psi_a(x) M_b = M_b
        worst = max(worst, np.abs(s - t[:6]).max(), np.abs(m - t[6:]).max())
    print(f"[sur] max |ratio(GPU sample) - ratio(oracle)| = {worst:.2e}", flush=True)
    assert worst < 2e-4  # features are FP32: the re-derived geometry differs from the sampled one by ~1e-7 relative
    assert np.all(T > -1e-6) and np.all(T < 1 + 1e-6)


def test_tol_zero_falls_back_exactly():
    mesh, kernel, conf = _case(n=24, delta=0.2, seed=4)
    ref = oracle.par_system(mesh, kernel, conf)
    A, st = _assemble(mesh, kernel, conf, sur=dict(tol=0.0))
    print(f"[sur] tol=0: cut {int(st['cut'])} taken {int(st['sur_pairs'])} handed back {int(st['sur_rejected'])}",
          flush=True)
    assert st["sur_rejected"] > 0 and st["sur_pairs"] + st["sur_rejected"] == st["cut"]
    assert np.array_equal(A.indptr, ref["A"].indptr) and np.array_equal(A.indices, ref["A"].indices)
    Abs = oracle.abs_matrix(mesh, kernel, conf).tocsr()
    D = (A - ref["A"]).tocoo()
    bound = 1e-12 * np.asarray(Abs[D.row, D.col]).ravel()
    assert np.all(np.abs(D.data) <= bound)


def test_default_model_accuracy_and_gain():
    """The trained ensemble against the exact matrix on an unseen mesh, and what training and the
    surrogate buy: untrained networks of the same shape, and the baryCenter method (the cheap
    method without learning), are both far less accurate."""
    mesh, kernel, conf = _case(n=64, delta=0.2, jitter=0.25, seed=2024)
    ref = oracle.par_system(mesh, kernel, conf)["A"]
    A, st = _assemble(mesh, kernel, conf, sur=dict())
    rel = sp.linalg.norm(A - ref) / sp.linalg.norm(ref)
    one = np.abs(A @ np.ones(A.shape[0])).max() / abs(A).max()
    acc = st["sur_pairs"] / st["cut"]
    print(f"[sur] default model: rel_fro {rel:.3e}, |A 1|/max|A| {one:.1e}, accepted {acc:.4f}", flush=True)
    assert np.array_equal(A.indices, ref.indices) and np.array_equal(A.indptr, ref.indptr)
    assert rel < 2e-5
    assert one < 1e-12
    assert acc > 0.95
    # untrained networks of the same architecture (and a permissive tol, so they are used)
    import tempfile, os
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "untrained.npz")
        _, norm, meta = surrogate.load()
        surrogate.save(path, np.stack([surrogate.init_params(1), surrogate.init_params(2)]), norm,
                       smin=meta["smin"], smax=meta["smax"], tol=1e9)
        B, stb = _assemble(mesh, kernel, conf, sur=dict(model=path))
    rel_u = sp.linalg.norm(B - ref) / sp.linalg.norm(ref)
    bc = make_configuration(method="baryCenter")
    C, _ = _assemble(mesh, kernel, bc)
    common = C.multiply(abs(ref) > 0)  # compare on the exact pattern
    rel_b = sp.linalg.norm(common - ref) / sp.linalg.norm(ref)
    print(f"[sur] rel_fro: trained {rel:.2e}, untrained {rel_u:.2e}, baryCenter {rel_b:.2e}", flush=True)
    assert rel < 0.01 * rel_u
    assert rel < 0.01 * rel_b
