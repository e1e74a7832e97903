"""Cut pairs of the oracle (integrate_retriangulate, src/integration.cpp:345-540) against an
independent quadrature of the same polygons.

For a kernel that is a polynomial in y (constant, parabola), every inner integrand of the
four blocks is a polynomial of degree <= 4 on the polygon B(x_k, delta) n bT, so the
reference's 7-point rule on its fan triangles (exact to degree 5, to the 14-digit rounding of
its literals) equals the exact integral over the polygon.  Here that integral is recomputed
from the polygon's corner list alone, by a different triangulation (a fan from the corners'
mean instead of from the first corner) and a different rule (8 x 8 Gauss-Legendre points
collapsed onto each triangle, exact to degree 14), with the basis of bT from numpy's solver.
A wrong fan, a wrong inner point map, a dropped rTdet or a block mixed up with another fails
at the 1e-12 level.  The polygon corners come from the oracle's method_retriangulate, which
tests/test_oracle.py pins by hand cases and disc/triangle properties; kernel point values come
from the oracle's kernels, pinned there by their closed-form moments.

The outer sums follow the definitions (src/integration.cpp:508-537), with w_k = dx_k |det aT|:
  termLocal[a][b]       = sum_k w_k psi_a(x_k) psi_b(x_k) int_P k(x_k, y)
  termNonloc[a][b]      = sum_k w_k psi_a(x_k) int_P psi_b(y) k(x_k, y)
  termNonlocPrime[a][b] = sum_k w_k psi_a(x_k) int_P psi_b(y) k(y, x_k)
  termLocalPrime[a][b]  = sum_k w_k int_P psi_a(y) psi_b(y) k(y, x_k)
"""
import numpy as np
import pytest

import oracle
from nlfem_amd.mesh import unit_square
from nlfem_amd.quadrature import P7, W7


def _collapsed_rule(m=8):
    """Gauss-Legendre on [0,1]^2 mapped to the reference triangle by (u, v) -> (u (1-v), v)."""
    g, w = np.polynomial.legendre.leggauss(m)
    t, wt = 0.5 * (g + 1), 0.5 * w
    U, V = np.meshgrid(t, t, indexing="ij")
    W = np.outer(wt, wt) * (1 - V)
    return np.stack([(U * (1 - V)).ravel(), V.ravel()], 1), W.ravel()


QP, QW = _collapsed_rule()


def _polygon(n, tris):
    """Corners of the fan (R0, R1, R2), (R0, R2, R3), ... returned by method_retriangulate."""
    T = np.asarray(tris[:6 * n]).reshape(n, 3, 2)
    return np.array([T[0, 0], T[0, 1]] + [T[j, 2] for j in range(n)])


def _points_weights(poly):
    """Quadrature points and weights (area measure) over a convex polygon, fan from its mean."""
    c = poly.mean(0)
    Y, Wt = [], []
    for j in range(len(poly)):
        p, q = poly[j], poly[(j + 1) % len(poly)]
        J = abs((p[0] - c[0]) * (q[1] - c[1]) - (q[0] - c[0]) * (p[1] - c[1]))
        Y.append(c + QP[:, :1] * (p - c) + QP[:, 1:] * (q - c))
        Wt.append(J * QW)
    return np.concatenate(Y), np.concatenate(Wt)


def _basis(E, Y):
    E = np.asarray(E).reshape(3, 2)
    M = np.column_stack([E[1] - E[0], E[2] - E[0]])
    p = np.linalg.solve(M, (Y - E[0]).T).T
    return np.column_stack([1 - p[:, 0] - p[:, 1], p[:, 0], p[:, 1]])


def _area2(E):
    E = np.asarray(E).reshape(3, 2)
    return abs((E[1, 0] - E[0, 0]) * (E[2, 1] - E[0, 1]) - (E[2, 0] - E[0, 0]) * (E[1, 1] - E[0, 1]))


def _independent_blocks(EA, EB, sqdelta, kname, caps, s=-1.0, od=1):
    """The four blocks from their definitions.  For outdim 2 the degree of freedom (node i,
    component c) is row/column 2 i + c, and the kernel's 2 x 2 matrix k_cd multiplies
    component d of the column into component c of the row."""
    n3 = 3 * od
    tL, tN, tLP, tNP = (np.zeros((n3, n3)) for _ in range(4))
    detA = _area2(EA)
    EAr = np.asarray(EA).reshape(3, 2)
    for k in range(P7.shape[0]):
        x = EAr[0] + P7[k, 0] * (EAr[1] - EAr[0]) + P7[k, 1] * (EAr[2] - EAr[0])
        n, tris = oracle.retriangulate(x, EB, sqdelta, caps)
        if n == 0:
            continue
        poly = _polygon(n, tris)
        Y, Wt = _points_weights(poly)
        # kernel matrices [npts, od, od] (row-major k00, k01, k10, k11 for outdim 2)
        kxy = np.array([oracle.kernel_value(kname, x, y, sqdelta, s)[:od * od] for y in Y]).reshape(-1, od, od)
        kyx = np.array([oracle.kernel_value(kname, y, x, sqdelta, s)[:od * od] for y in Y]).reshape(-1, od, od)
        psiY = _basis(EB, Y)
        psix = np.array([1 - P7[k, 0] - P7[k, 1], P7[k, 0], P7[k, 1]])
        wk = W7[k] * detA
        IL = np.einsum("q,qcd->cd", Wt, kxy)             # int_P k(x, y)
        IN = np.einsum("q,qj,qcd->jcd", Wt, psiY, kxy)    # int_P psi_j(y) k(x, y)
        INP = np.einsum("q,qj,qcd->jcd", Wt, psiY, kyx)   # int_P psi_j(y) k(y, x)
        ILP = np.einsum("q,qi,qj,qcd->ijcd", Wt, psiY, psiY, kyx)  # int_P psi_i psi_j k(y, x)
        for i in range(3):
            for c in range(od):
                for j in range(3):
                    for d in range(od):
                        r, col = i * od + c, j * od + d
                        tL[r, col] += wk * psix[i] * psix[j] * IL[c, d]
                        tN[r, col] += wk * psix[i] * IN[j, c, d]
                        tNP[r, col] += wk * psix[i] * INP[j, c, d]
                        tLP[r, col] += wk * ILP[i, j, c, d]
    return tL, tN, tLP, tNP


def _cut_pairs(mesh, delta, count, seed):
    """Pairs (aT, bT) of the oracle's neighbour lists in which some outer point's polygon is a
    proper part of bT (or empty): the cut pairs."""
    conf = {"quadrature": {"outer": {"points": P7, "weights": W7}, "inner": {"points": P7, "weights": W7}}}
    from nlfem_amd.config import make_configuration
    r = oracle.par_system(mesh, {"function": "constant", "horizon": delta, "outputdim": 1}, make_configuration(),
                          keep_neighbours=True)
    ptr, nbr = r["nbr_ptr"], r["nbr"]
    el, vt = np.asarray(mesh["elements"]), np.asarray(mesh["vertices"])
    rng = np.random.default_rng(seed)
    out = []
    for aT in rng.permutation(len(ptr) - 1):
        if ptr[aT + 1] == ptr[aT]:
            continue
        EA = vt[el[aT]].ravel()
        for bT in rng.permutation(nbr[ptr[aT]:ptr[aT + 1]])[:6]:
            EB = vt[el[bT]].ravel()
            EAr = EA.reshape(3, 2)
            full = True
            for k in range(P7.shape[0]):
                x = EAr[0] + P7[k, 0] * (EAr[1] - EAr[0]) + P7[k, 1] * (EAr[2] - EAr[0])
                n, tris = oracle.retriangulate(x, EB, delta * delta, 1)
                a = 0.0 if n == 0 else sum(_area2(tris[6 * j:6 * j + 6]) for j in range(n))
                full = full and abs(a - _area2(EB)) <= 1e-12 * _area2(EB)
            if not full:
                out.append((int(aT), int(bT)))
                break
        if len(out) >= count:
            return out
    return out


@pytest.mark.parametrize("kname", ["constant", "parabola"])
def test_cut_pair_blocks_match_independent_quadrature(kname):
    delta = 0.2
    mesh = unit_square(12, delta, jitter=0.2, seed=11)
    pairs = _cut_pairs(mesh, delta, 12, seed=3)
    assert len(pairs) >= 10
    el, vt = np.asarray(mesh["elements"]), np.asarray(mesh["vertices"])
    for aT, bT in pairs:
        r, (tL, tN, tLP, tNP) = oracle.integrate_pair("retriangulate", el, vt, aT, bT, P7, W7, P7, W7,
                                                     delta * delta, kname)
        assert r
        ref = _independent_blocks(vt[el[aT]].ravel(), vt[el[bT]].ravel(), delta * delta, kname, 1)
        scale = max(np.abs(b).max() for b in ref)
        for name, got, want in zip(("termLocal", "termNonloc", "termLocalPrime", "termNonlocPrime"),
                                   (tL, tN, tLP, tNP), ref):
            err = np.abs(got - want).max() / scale
            assert err < 1e-12, f"{kname} pair {aT},{bT} {name}: {err:.2e}"


def test_independent_quadrature_detects_planted_mistakes():
    """The check above fails for a plausible mistake: the basis of the outer element used on
    the inner side, a fan missing its last triangle, or the outer weights of another point."""
    delta = 0.2
    mesh = unit_square(12, delta, jitter=0.2, seed=11)
    aT, bT = _cut_pairs(mesh, delta, 1, seed=3)[0]
    el, vt = np.asarray(mesh["elements"]), np.asarray(mesh["vertices"])
    EA, EB = vt[el[aT]].ravel(), vt[el[bT]].ravel()
    _, blocks = oracle.integrate_pair("retriangulate", el, vt, aT, bT, P7, W7, P7, W7, delta * delta, "parabola")
    ref = _independent_blocks(EA, EB, delta * delta, "parabola", 1)
    # termNonloc with the basis of aT instead of bT inside the polygon
    bad = _independent_blocks(EA, EA, delta * delta, "parabola", 1)
    assert np.abs(bad[1] - blocks[1]).max() > 1e-6 * np.abs(ref[1]).max()
    # a fan without its last triangle
    global _polygon
    keep = _polygon
    try:
        _polygon = lambda n, tris: keep(max(n - 1, 1), tris)
        bad = _independent_blocks(EA, EB, delta * delta, "parabola", 1)
    finally:
        _polygon = keep
    assert max(np.abs(b - g).max() for b, g in zip(bad, blocks)) > 1e-6 * np.abs(ref[1]).max()
    # termLocal and termLocalPrime swapped
    assert np.abs(blocks[0] - ref[2]).max() > 1e-6 * np.abs(ref[0]).max()


def _separation(EA, EB):
    A = np.asarray(EA).reshape(3, 2)
    B = np.asarray(EB).reshape(3, 2)
    return min(np.linalg.norm(a - b) for a in A for b in B)


@pytest.mark.parametrize("kname,s", [("fractional", 0.4), ("linearPrototypeMicroelastic", -0.5),
                                     ("linearPrototypeMicroelasticField", -0.5)])
def test_cut_pairs_of_non_polynomial_kernels_consistent(kname, s):
    """For the non-polynomial kernels the reference's 7-point fan sums are not exact integrals,
    so they have no exact value to be pinned against at 1e-12.  On cut pairs whose elements lie
    at least two mesh sizes apart (the kernel is smooth on the scale of the triangles) they must
    still agree with the independent high-order polygon quadrature to the 7-point rule's
    truncation error: a kernel evaluated at the wrong points, or with the wrong exponent, fails."""
    delta, n = 0.2, 12
    mesh = unit_square(n, delta, jitter=0.2, seed=11)
    el, vt = np.asarray(mesh["elements"]), np.asarray(mesh["vertices"])
    pairs = [(a, b) for a, b in _cut_pairs(mesh, delta, 60, seed=5)
             if _separation(vt[el[a]], vt[el[b]]) >= 2.0 / n][:12]
    assert len(pairs) >= 8
    od = 2 if kname.endswith("Field") else 1
    worst = 0.0
    for aT, bT in pairs:
        r, got = oracle.integrate_pair("retriangulate", el, vt, aT, bT, P7, W7, P7, W7, delta * delta, kname,
                                       fractional_s=s, outdim=od)
        ref = _independent_blocks(vt[el[aT]].ravel(), vt[el[bT]].ravel(), delta * delta, kname, 1, s, od)
        scale = max(np.abs(b).max() for b in ref)
        worst = max(worst, max(np.abs(g - w).max() / scale for g, w in zip(got, ref)))
    print(f"[cut] {kname}: separated cut pairs, 7-point sums vs high-order quadrature {worst:.2e}")
    assert worst < 1e-4
    if od == 2:
        # component layout mistaken (the 2 x 2 kernel blocks transposed into the node blocks, i.e.
        # row 3 c + i instead of 2 i + c) and the kernel's diagonal exchanged are far outside
        aT, bT = pairs[0]
        _, got = oracle.integrate_pair("retriangulate", el, vt, aT, bT, P7, W7, P7, W7, delta * delta, kname,
                                       fractional_s=s, outdim=2)
        ref = _independent_blocks(vt[el[aT]].ravel(), vt[el[bT]].ravel(), delta * delta, kname, 1, s, 2)
        perm = np.array([0, 2, 4, 1, 3, 5])
        scale = max(np.abs(b).max() for b in ref)
        assert max(np.abs(g[np.ix_(perm, perm)] - w).max() for g, w in zip(got, ref)) > 1e-2 * scale
        with oracle.planted("field_diagonal"):
            _, bad = oracle.integrate_pair("retriangulate", el, vt, aT, bT, P7, W7, P7, W7, delta * delta, kname,
                                           fractional_s=s, outdim=2)
        assert max(np.abs(g - w).max() for g, w in zip(bad, ref)) > 1e-2 * scale
    # the exponent of the fractional kernel mistaken (2 + s instead of 2 + 2s) is far outside
    if kname == "fractional":
        aT, bT = pairs[0]
        _, got = oracle.integrate_pair("retriangulate", el, vt, aT, bT, P7, W7, P7, W7, delta * delta, kname,
                                       fractional_s=s)
        bad = _independent_blocks(vt[el[aT]].ravel(), vt[el[bT]].ravel(), delta * delta, kname, 1, 0.2)
        assert max(np.abs(g - w).max() for g, w in zip(got, bad)) > 1e-2 * max(np.abs(b).max() for b in bad)
