"""Pinning the CPU oracle (oracle/nlfem_oracle.cpp) against what the mathematics fixes.

The reference holds no golden vectors for this path (its tests cover an unused mesh builder)
and cannot be built here (Armadillo, METIS and CGAL are absent), so each oracle function is
checked against closed forms, invariants, brute force and independent quadrature, chosen so a
dropped term, a wrong sign or index, or a transposed operand fails at least one check.
"""
import itertools
import math

import numpy as np
import pytest
import scipy.integrate as si
import scipy.sparse as sp

import oracle
from helpers import case, fan_mesh, gmsh_case, label0_mesh
from nlfem_amd.config import make_configuration
from nlfem_amd.mesh import mesh_dict, unit_square
from nlfem_amd.quadrature import P7, W7, P16, W16, tensorgauss


# ------------------------------------------------------------------ quadrature literals
def _tri_monomial(a, b):
    """Exact integral of x^a y^b over the reference triangle: a! b! / (a+b+2)!."""
    return math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)


def test_seven_point_rule_literals():
    assert W7.sum() == pytest.approx(0.5, abs=2e-14)
    for a in range(6):
        for b in range(6 - a):
            q = np.sum(W7 * P7[:, 0] ** a * P7[:, 1] ** b)
            assert q == pytest.approx(_tri_monomial(a, b), rel=1e-12, abs=1e-14)
    # chk_BasisFunction (src/checks.cpp:36-47): integral of psi_0 = 1/6
    assert abs(np.sum(W7 * (1 - P7[:, 0] - P7[:, 1])) - 1 / 6) < 1e-8
    assert abs(np.sum(W16 * (1 - P16[:, 0] - P16[:, 1])) - 1 / 6) < 1e-8


def test_tensorgauss_layout():
    tg = tensorgauss(3)
    assert tg.points.shape == (81, 4)
    assert tg.weights.sum() == pytest.approx(16.0, rel=1e-14)
    # 'ij' order, last index fastest
    assert np.array_equal(tg.points[0, :3], tg.points[1, :3]) and tg.points[0, 3] != tg.points[1, 3]
    p, w = np.polynomial.legendre.leggauss(3)
    assert np.array_equal(tg.points[5], [p[0], p[0], p[1], p[2]])


# ------------------------------------------------------------------ kernels
@pytest.mark.parametrize("name,s", [("constant", -1.0), ("parabola", -1.0), ("fractional", 0.4),
                                    ("fractional", 0.1), ("linearPrototypeMicroelastic", -0.5)])
def test_kernel_second_moment_is_two(name, s):
    """Every scalar kernel is normalised so that int_{B_delta} |z|^2 k(z) dz = 2."""
    delta = 0.13
    sq = delta ** 2
    f = lambda r: oracle.kernel_value(name, [0.3 + r, 0.2], [0.3, 0.2], sq, s)[0] * r ** 3 * 2 * np.pi
    val, _ = si.quad(f, 0, delta, limit=200)
    assert val == pytest.approx(2.0, rel=1e-8)


def test_kernel_field_trace_moment_and_shape():
    delta = 0.17
    sq = delta ** 2
    z = np.array([0.03, -0.05])
    k = oracle.kernel_value("linearPrototypeMicroelasticField", z, [0, 0], sq, -0.5)
    K = k.reshape(2, 2)
    assert np.allclose(K, K.T, rtol=0, atol=1e-12 * abs(K).max())
    # rank one along z: K z_perp = 0
    assert abs(K @ np.array([-z[1], z[0]])).max() < 1e-12 * abs(K).max()
    f = lambda r: np.trace(oracle.kernel_value("linearPrototypeMicroelasticField", [r, 0], [0, 0], sq, -0.5)
                           .reshape(2, 2)) * r ** 3 * 2 * np.pi
    val, _ = si.quad(f, 0, delta)
    assert val == pytest.approx(2 * np.pi, rel=1e-10)  # (3/delta^3) * 2*pi*delta^3/3


def test_kernels_symmetric_in_arguments():
    rng = np.random.default_rng(0)
    for name, s in [("parabola", -1), ("fractional", 0.4), ("linearPrototypeMicroelasticField", -0.5)]:
        for _ in range(20):
            x, y = rng.uniform(size=2), rng.uniform(size=2)
            assert np.array_equal(oracle.kernel_value(name, x, y, 0.04, s), oracle.kernel_value(name, y, x, 0.04, s))


# ------------------------------------------------------------------ geometry helpers
def test_toref_inverts_tophys():
    rng = np.random.default_rng(1)
    for _ in range(200):
        E = rng.uniform(size=6)
        if abs((E[2] - E[0]) * (E[5] - E[1]) - (E[4] - E[0]) * (E[3] - E[1])) < 1e-3:
            continue
        p = rng.uniform(size=2)
        assert np.allclose(oracle.to_ref(E, oracle.to_phys(E, p)), p, atol=1e-10)


def test_solve2x2_pivoting_and_singular():
    A = np.array([1e-3, 2.0, 5.0, 1.0])  # row-major [[a00,a01],[a10,a11]]
    b = np.array([1.0, 2.0])
    x = oracle.solve2x2(A, b)
    assert np.allclose(A.reshape(2, 2) @ x, b, atol=1e-13)
    with pytest.raises(ValueError):
        oracle.solve2x2([1.0, 2.0, 2.0, 4.0], b)


# ------------------------------------------------------------------ retriangulation
def _poly_area(lst, n):
    a = 0.0
    for t in range(n):
        T = lst[6 * t:6 * t + 6]
        a += 0.5 * abs((T[2] - T[0]) * (T[5] - T[1]) - (T[4] - T[0]) * (T[3] - T[1]))
    return a


def test_retriangulate_hand_cases():
    T = np.array([0.0, 0.0, 1.0, 0.0, 0.0, 1.0])
    n, _ = oracle.retriangulate([5.0, 5.0], T, 0.25)
    assert n == 0  # no intersection
    n, l = oracle.retriangulate([0.3, 0.3], T, 4.0)
    assert n == 1 and np.array_equal(l[:6], T)  # all vertices inside: T itself, vertex order kept
    n, _ = oracle.retriangulate([0.25, 0.25], T, 0.01)
    assert n == 0  # ball strictly inside the triangle: no vertex, no edge point (reference quirk)
    # one edge cut twice, caps on: edge points + cap point -> 3 points -> 1 triangle
    n, l = oracle.retriangulate([0.5, -0.05], T, 0.1 ** 2, caps=1)
    assert n == 1
    pts = l[:6].reshape(3, 2)
    assert np.allclose(np.hypot(pts[:, 0] - 0.5, pts[:, 1] + 0.05), 0.1, atol=1e-14)
    n0, _ = oracle.retriangulate([0.5, -0.05], T, 0.1 ** 2, caps=0)
    assert n0 == 0  # without caps only the two edge points remain


def _disk_triangle_area_mc(x, T, delta, rng, n=400000):
    P = rng.uniform(size=(n, 2))
    P = P[P.sum(axis=1) <= 1.0]  # reference triangle samples
    E = T.reshape(3, 2)
    Y = E[0] + P[:, :1] * (E[1] - E[0]) + P[:, 1:] * (E[2] - E[0])
    area = 0.5 * abs((E[1, 0] - E[0, 0]) * (E[2, 1] - E[0, 1]) - (E[2, 0] - E[0, 0]) * (E[1, 1] - E[0, 1]))
    inside = np.hypot(Y[:, 0] - x[0], Y[:, 1] - x[1]) <= delta
    return area * inside.mean(), area / np.sqrt(len(P))


def test_retriangulate_polygon_inside_disk_and_triangle():
    rng = np.random.default_rng(3)
    checked = 0
    for _ in range(300):
        T = rng.uniform(size=6)
        E = T.reshape(3, 2)
        det = (E[1, 0] - E[0, 0]) * (E[2, 1] - E[0, 1]) - (E[2, 0] - E[0, 0]) * (E[1, 1] - E[0, 1])
        if abs(det) < 0.05:
            continue
        x = rng.uniform(-0.2, 1.2, size=2)
        delta = rng.uniform(0.1, 0.6)
        n, l = oracle.retriangulate(x, T, delta ** 2, caps=1)
        if n == 0:
            continue
        pts = l[:6 * n].reshape(-1, 2)
        assert np.all(np.hypot(pts[:, 0] - x[0], pts[:, 1] - x[1]) <= delta * (1 + 1e-12))
        M = np.array([[E[1, 0] - E[0, 0], E[2, 0] - E[0, 0]], [E[1, 1] - E[0, 1], E[2, 1] - E[0, 1]]])
        bc = np.linalg.solve(M, (pts - E[0]).T)
        assert np.all(bc >= -1e-10) and np.all(bc.sum(axis=0) <= 1 + 1e-10)
        # polygon is inscribed in disk n triangle: its area cannot exceed the exact area
        exact, err = _disk_triangle_area_mc(x, T, delta, rng, 100000)
        poly = _poly_area(l, n)
        assert poly <= exact + 5 * err + 1e-12
        checked += 1
    assert checked > 50


# ------------------------------------------------------------------ fully contained pairs
def test_fully_contained_pair_closed_form():
    """Constant kernel, inner element inside every ball: the 7-point rule is exact for the
    degree <= 2 integrands, so the four blocks have closed forms in |det|."""
    m = unit_square(8, 0.4)
    el, vt = m["elements"], m["vertices"]
    aT, bT = 18, 21
    delta = 0.6
    c = 4 / (np.pi * delta ** 4)
    r, (tL, tN, tLP, tNP) = oracle.integrate_pair("retriangulate", el, vt, aT, bT, P7, W7, P7, W7, delta ** 2,
                                                  "constant")
    dA = 2 * 0.5 / 64
    dB = dA
    mass = (np.ones((3, 3)) + np.eye(3)) / 24.0  # int psi_a psi_b on the reference triangle
    assert r == 1
    assert np.allclose(tN, c * dA * dB / 36.0, rtol=1e-12)
    assert np.allclose(tNP, tN, rtol=0, atol=0)
    assert np.allclose(tL, c * dA * dB * mass * 0.5, rtol=1e-12)
    assert np.allclose(tLP, c * dA * dB * mass * 0.5, rtol=1e-12)


# ------------------------------------------------------------------ Duffy tables
@pytest.mark.parametrize("case_id", [0, 1, 2])
def test_duffy_tables_measure_and_polynomials(case_id):
    """With s = 0 the canceled determinant is the full Jacobian: the rule integrates over
    T x T (measure 1/4) and is exact for low-degree polynomials of (x, y)."""
    al, ac, df, dw, dg = oracle.quadrule_tables(5, 0.0, case_id)
    w = df * dg[None, :]
    assert w.sum() == pytest.approx(0.25, rel=1e-13)
    x0, x1, y0, y1 = al[..., 0], al[..., 1], al[..., 2], al[..., 3]
    assert np.all(x0 >= -1e-14) and np.all(x0 + x1 <= 1 + 1e-14) and np.all(y0 + y1 <= 1 + 1e-14)
    assert np.sum(w * x0 * y1) == pytest.approx(1 / 36, rel=1e-12)
    assert np.sum(w * (x0 - y0) ** 2) == pytest.approx(1 / 36, rel=1e-12)
    assert np.sum(w * x1 ** 2 * y0) == pytest.approx((1 / 12) * (1 / 6), rel=1e-12)
    # canceled points have xi = 1 before the mirror: alphaC = alpha / xi
    assert np.all(ac[..., 0] + ac[..., 1] >= 1 - 1e-12) or case_id >= 0


def _pair_form(E_list, shared_nodes, c, nq=P7, wq=W7):
    """c * int_X int_Y (phi_i(x)-phi_i(y)) (phi_j(x)-phi_j(y)) dx dy summed over the ordered
    element pairs (X, Y) in E_list, by an independent 7x7 product rule (exact for degree 2)."""
    nodes = sorted(set(n for X in shared_nodes for n in X))
    idx = {n: i for i, n in enumerate(nodes)}
    M = np.zeros((len(nodes), len(nodes)))
    for (EX, NX), (EY, NY) in E_list:
        dX = abs(np.linalg.det(np.array([EX[1] - EX[0], EX[2] - EX[0]])))
        dY = abs(np.linalg.det(np.array([EY[1] - EY[0], EY[2] - EY[0]])))
        psi = np.stack([1 - nq[:, 0] - nq[:, 1], nq[:, 0], nq[:, 1]], axis=1)  # [q, 3]
        for i in range(7):
            for j in range(7):
                u = np.zeros(len(nodes))
                for a in range(3):
                    u[idx[NX[a]]] += psi[i, a]
                for b in range(3):
                    u[idx[NY[b]]] -= psi[j, b]
                M += c * wq[i] * wq[j] * dX * dY * np.outer(u, u)
    return nodes, M


@pytest.mark.parametrize("method", ["fractional", "weakSingular"])
@pytest.mark.parametrize("layout", ["vertex", "edge"])
def test_close_pair_routines_at_s_minus_one(method, layout):
    """fractional_s = -1 makes the fractional kernel the constant c = 4/(pi delta^4) and the
    canceled determinant the full Jacobian; both close-pair routines then assemble
    c * int int (u(x)-u(y))(v(x)-v(y)) over every ordered pair, which the test integrates
    independently (argument sorting, signs and the block scatter are all exercised)."""
    if layout == "vertex":
        verts = np.array([[0.0, 0.0], [0.1, 0.0], [0.0, 0.1], [-0.1, 0.02], [-0.03, -0.1]])
        elems = np.array([[0, 1, 2], [0, 3, 4]])
    else:
        verts = np.array([[0.0, 0.0], [0.1, 0.0], [0.0, 0.1], [0.12, 0.09]])
        elems = np.array([[0, 1, 2], [2, 1, 3]])
    nV = len(verts)
    mesh = mesh_dict(elems, np.array([1, 1]), verts, np.ones(nV, dtype=np.int64))
    delta = 0.3
    kern = {"function": "fractional", "horizon": delta, "outputdim": 1, "fractional_s": -1.0}
    conf = make_configuration(close=method, tensorGaussDegree=4)
    r = oracle.par_system(mesh, kern, conf)
    A = r["A"].toarray()
    c = 4 / (np.pi * delta ** 4)
    E = [(verts[e], list(e)) for e in elems]
    pairs = [(E[i], E[j]) for i in range(2) for j in range(2)]
    nodes, M = _pair_form(pairs, [list(e) for e in elems], c)
    assert r["n_close"] == 4
    assert np.allclose(A[np.ix_(nodes, nodes)], M, rtol=1e-10, atol=1e-12 * abs(M).max())


# ------------------------------------------------------------------ dual graph and BFS
def test_dual_graph_brute_force():
    m = unit_square(7, 0.1, jitter=0.2, seed=4)
    el = m["elements"]
    xadj, adj = oracle.dual_graph(el, m["nV"])
    for e in range(len(el)):
        brute = [f for f in range(len(el)) if f != e and len(set(el[e]) & set(el[f])) > 0]
        assert list(adj[xadj[e]:xadj[e + 1]]) == brute
    # an interior triangle of a regular mesh has exactly 12 vertex-sharing neighbours
    xadj, adj = oracle.dual_graph(unit_square(6, 0.1)["elements"], 49)
    assert max(np.diff(xadj)) == 12


def _interacts(mesh, conf, kernel, a, b):
    q = conf["quadrature"]
    r, _ = oracle.integrate_pair(conf["approxBalls"]["method"], mesh["elements"], mesh["vertices"], a, b,
                                 q["outer"]["points"], q["outer"]["weights"], q["inner"]["points"],
                                 q["inner"]["weights"], kernel["horizon"] ** 2, kernel["function"],
                                 caps=conf["approxBalls"].get("is_PlacePointOnCap", 1))
    return r != 0


@pytest.mark.parametrize("which", ["C1", "label0", "bary", "hole"])
def test_bfs_equals_connected_component(which):
    """Neighbour list of aT = elements that interact with aT and are reachable from aT in the
    vertex-sharing dual graph through interacting (or label-0) elements."""
    if which == "label0":
        mesh, kernel, conf = label0_mesh()
    elif which == "hole":  # the tutorial mesh with its artificial vertex (128 label-0 elements)
        mesh, kernel, conf = gmsh_case("hole")
    elif which == "bary":
        mesh, kernel, conf = case("bary")
        mesh = unit_square(8, 0.15, jitter=0.1, seed=1)
    else:
        mesh, kernel, conf = case("C1")
        mesh = unit_square(9, 0.1)
    res = oracle.par_system(mesh, kernel, conf, keep_neighbours=True)
    el = mesh["elements"]
    lab = mesh["elementLabels"]
    nE = len(el)
    xadj, adj = oracle.dual_graph(el, mesh["nV"])
    rng = np.random.default_rng(0)
    for aT in rng.choice(np.where(lab != 0)[0], size=min(25, nE), replace=False):
        inter = {b for b in range(nE) if lab[b] != 0 and _interacts(mesh, conf, kernel, aT, b)}
        passable = inter | {b for b in range(nE) if lab[b] == 0}
        seen, stack, comp = {aT}, [aT], set()
        if aT in inter:
            comp.add(aT)
        while stack:
            s = stack.pop()
            for b in adj[xadj[s]:xadj[s + 1]]:
                if b in passable and b not in seen:
                    seen.add(b)
                    stack.append(b)
                    if b in inter:
                        comp.add(b)
        got = set(res["nbr"][res["nbr_ptr"][aT]:res["nbr_ptr"][aT + 1]].tolist())
        assert got == comp, aT


# ------------------------------------------------------------------ assembled matrices
@pytest.mark.parametrize("name", ["C1", "C2s", "C3s", "C5s", "DG", "bary", "weak", "nocaps", "hole", "disjoint",
                                  "circle", "fan"])
def test_matrix_invariants(name):
    """A 1 = 0 on every assembled row (partition of unity in all four blocks and in the
    basis differences), and the Omega x Omega block is symmetric for these kernels."""
    if name == "fan":
        mesh, kernel, conf = fan_mesh(140)
    else:
        mesh, kernel, conf = gmsh_case(name) if name in ("hole", "disjoint", "circle") else case(name)
    r = oracle.par_system(mesh, kernel, conf)
    A = r["A"]
    rows = np.where(mesh["nodeLabels"] > 0)[0]
    scale = abs(A).max()
    tol = 1e-8 if name in ("C5s", "circle") else 1e-13  # C5s: near-coincident outer/inner points on identical pairs
    assert np.abs(np.asarray(A[rows].sum(axis=1))).max() <= tol * scale
    if name not in ("DG", "bary"):  # baryCenter pairs integrate one direction only (factor 2, no Prime blocks)
        AO = A[rows][:, rows].toarray()
        assert np.abs(AO - AO.T).max() <= (1e-8 if name in ("C5s", "circle") else 1e-13) * scale


def test_thread_count_independent():
    mesh, kernel, conf = case("C2s")
    A1 = oracle.par_system(mesh, kernel, conf, nthreads=1)["A"]
    A8 = oracle.par_system(mesh, kernel, conf, nthreads=8)["A"]
    assert np.array_equal(A1.indptr, A8.indptr) and np.array_equal(A1.indices, A8.indices)
    assert abs(A1 - A8).max() <= 1e-14 * abs(A1).max()


def test_consistency_with_laplacian_for_quadratics():
    """For u = x0^2 x1 + x1^2 (-Laplace u = -2(x1+1)) the rows of A u approach the load vector
    of f = -2(x1+1) as h -> 0 (the manufactured solution of the reference's rates harness)."""
    errs = []
    for n in (16, 32):
        mesh = unit_square(n, 0.1)
        kern = {"function": "constant", "horizon": 0.1, "outputdim": 1}
        conf = make_configuration()
        A = oracle.par_system(mesh, kern, conf)["A"]
        x = mesh["vertices"]
        u = x[:, 0] ** 2 * x[:, 1] + x[:, 1] ** 2
        f = oracle.load_vector(mesh, {"function": "linear"}, conf)
        rows = np.where(mesh["vertexLabels"] > 0)[0]
        errs.append(np.linalg.norm((A @ u - f)[rows]) / np.linalg.norm(f[rows]))
    assert errs[1] < errs[0] and errs[1] < 0.05
