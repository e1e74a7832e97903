"""Domain-decomposition weights (nZeta > 0) of the oracle.

The reference weighs every addend of pair (aT, bT) by 1 / |Z_aT n Z_bT|, the number of
subdomains the two elements share (evaluateZeta, src/mathhelpers.cpp:19-38, applied at
src/Cassemble.cpp:648-656 and :741-766), and every load contribution of aT by 1 / |Z_aT|
(:906-920).  Z is the CSR indicator mesh["ZetaIndicator"] (elements x subdomains) with
mesh["nZeta"] columns (cython/nlfem.pyx:205-214).

Pins: with every element in the same k subdomains the weighted matrix is A / k; for a random
cover the weighted matrix equals the sum over the oracle's own neighbour pairs of
w(aT, bT) times each pair's blocks, scattered by the reference's index rule in this file (the
same scatter with w = 1 reproduces the unweighted matrix, which the other oracle tests pin),
with w from Python set intersections; the weighted load vector follows from linearity.
"""
import numpy as np
import scipy.sparse as sp

import oracle
from nlfem_amd.config import make_configuration
from nlfem_amd.mesh import unit_square


def zeta_indicator(rows, ncols):
    """CSR indicator matrix (sorted rows) from a list of subdomain sets per element."""
    indptr = np.zeros(len(rows) + 1, dtype=np.int64)
    indices = []
    for e, r in enumerate(rows):
        s = sorted(r)
        indices += s
        indptr[e + 1] = indptr[e] + len(s)
    data = np.ones(len(indices))
    return sp.csr_matrix((data, np.asarray(indices, dtype=np.int64), indptr), shape=(len(rows), ncols))


def random_cover(nE, seed):
    """Every element in subdomain 0 (so every pair shares one), about half also in 1, a fifth in 2."""
    rng = np.random.default_rng(seed)
    rows = []
    for e in range(nE):
        r = {0}
        if rng.random() < 0.5:
            r.add(1)
        if rng.random() < 0.2:
            r.add(2)
        rows.append(r)
    return rows


def with_zeta(mesh, rows, ncols=3):
    m = dict(mesh)
    m["ZetaIndicator"] = zeta_indicator(rows, ncols)
    m["nZeta"] = ncols
    return m


def _case():
    mesh = unit_square(10, 0.15, jitter=0.2, seed=2)
    kern = {"function": "constant", "horizon": 0.15, "outputdim": 1}
    return mesh, kern, make_configuration()


def test_uniform_cover_scales_by_one_over_k():
    mesh, kern, conf = _case()
    A = oracle.par_system(mesh, kern, conf)["A"]
    B = oracle.par_system(with_zeta(mesh, [{0, 1, 2}] * mesh["nE"]), kern, conf)["A"]
    assert np.array_equal(A.indptr, B.indptr) and np.array_equal(A.indices, B.indices)
    assert np.abs(B.data - A.data / 3).max() <= 1e-15 * np.abs(A.data).max()
    f = oracle.load_vector(mesh, {"function": "linear"}, conf)
    fz = oracle.load_vector(with_zeta(mesh, [{0, 1, 2}] * mesh["nE"]), {"function": "linear"}, conf)
    assert np.allclose(fz, f / 3, rtol=1e-14, atol=0)


def _pairwise_matrix(mesh, kern, conf, weight):
    """sum over the oracle's neighbour pairs of weight(aT, bT) x the pair's four blocks,
    scattered as src/Cassemble.cpp:721-770 does (CG, outdim 1)."""
    r = oracle.par_system(mesh, kern, conf, keep_neighbours=True)
    el, vt, vl = mesh["elements"], mesh["vertices"], mesh["vertexLabels"]
    q = conf["quadrature"]
    K = mesh["K"]
    rows, cols, vals = [], [], []
    ptr, nbr = r["nbr_ptr"], r["nbr"]
    for aT in range(len(ptr) - 1):
        for bT in nbr[ptr[aT]:ptr[aT + 1]]:
            _, (tL, tN, tLP, tNP) = oracle.integrate_pair("retriangulate", el, vt, aT, bT, q["outer"]["points"],
                                                          q["outer"]["weights"], q["inner"]["points"],
                                                          q["inner"]["weights"], kern["horizon"] ** 2,
                                                          kern["function"])
            w = weight(aT, int(bT))
            aA, bA = el[aT], el[bT]
            for a in range(3):
                for b in range(3):
                    if vl[aA[a]] > 0:
                        rows += [aA[a], aA[a]]
                        cols += [aA[b], bA[b]]
                        vals += [w * tL[a, b], -w * tN[a, b]]
                    if vl[bA[b]] > 0:
                        rows += [bA[b], bA[b]]
                        cols += [bA[a], aA[a]]
                        vals += [w * tLP[a, b], -w * tNP[a, b]]
    return sp.csr_matrix((vals, (rows, cols)), shape=(K, K)), r["A"]


def test_random_cover_equals_weighted_pair_sum():
    mesh, kern, conf = _case()
    rows = random_cover(mesh["nE"], seed=4)
    # the scatter of this file reproduces the unweighted oracle
    P1, A = _pairwise_matrix(mesh, kern, conf, lambda a, b: 1.0)
    assert abs(P1 - A).max() <= 1e-13 * abs(A).max()
    Pz, _ = _pairwise_matrix(mesh, kern, conf, lambda a, b: 1.0 / len(rows[a] & rows[b]))
    Az = oracle.par_system(with_zeta(mesh, rows), kern, conf)["A"]
    assert abs(Pz - Az).max() <= 1e-13 * abs(Az).max()
    # the weights make a difference (pairs inside subdomain 1 count half, ...)
    assert abs(Az - A).max() > 0.1 * abs(A).max()


def test_planted_one_sided_weight_fails():
    """A weight taken from aT's row alone (1 / |Z_aT|) instead of the intersection differs."""
    mesh, kern, conf = _case()
    rows = random_cover(mesh["nE"], seed=4)
    Pbad, _ = _pairwise_matrix(mesh, kern, conf, lambda a, b: 1.0 / len(rows[a]))
    Az = oracle.par_system(with_zeta(mesh, rows), kern, conf)["A"]
    assert abs(Pbad - Az).max() > 1e-3 * abs(Az).max()


def test_load_vector_weights_by_linearity():
    """f_Z = f - 1/2 f_X - 1/6 f_Y ... in general f_Z = sum_e f_e / |Z_e|: with every element in
    subdomain 0 and X also in 1, f_Z = f - f_X / 2, where f_X is the load of X alone."""
    mesh, kern, conf = _case()
    rng = np.random.default_rng(8)
    X = rng.random(mesh["nE"]) < 0.5
    rows = [{0, 1} if x else {0} for x in X]
    f = oracle.load_vector(mesh, {"function": "linear"}, conf)
    fz = oracle.load_vector(with_zeta(mesh, rows, 2), {"function": "linear"}, conf)
    mx = dict(mesh)
    lab = mesh["elementLabels"].copy()
    lab[~X] = -1  # elements outside X contribute nothing (labels <= 0 are skipped)
    mx["elementLabels"] = lab
    fx = oracle.load_vector(mx, {"function": "linear"}, conf)
    assert np.allclose(fz, f - 0.5 * fx, rtol=1e-13, atol=1e-15 * np.abs(f).max())
