"""GPU parity: libnlfem_amd.so (HIP, through the C ABI) against the FP64 CPU oracle.

Bar: neighbour lists and sparsity pattern bit-exact; every FP64 entry within 1e-12 of the
oracle, relative to the sum of the absolute values of the addends that formed it (for
entries without cancellation that is the entry itself; the microelastic 16/7-point case has
near-coincident outer and inner points on identical pairs, whose addends exceed the entries
by ~1e6), plus 1e-15 of the largest entry of the matrix; mixed precision within 1e-5
relative Frobenius error.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from helpers import (case, check_floor_usage, compare, fan_mesh, floor_usage, gmsh_case, label0_mesh, rigid_residuals,
                     sorted_lists)
from nlfem_amd import nlfem
from nlfem_amd.config import make_configuration
from nlfem_amd.mesh import unit_square

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def _gpu(mesh, kernel, conf, paths=0):
    asm = nlfem.Assembler(mesh, kernel, conf)
    if paths:
        asm.force_paths(paths)
    try:
        asm.assemble()
        A = asm.csr()
        nb = asm.neighbours()
        st = asm.stats()
    finally:
        asm.close()
    return A, nb, st


def _check_case(name, mesh, kernel, conf, tol=FP64_TOL, paths=0):
    print(f"[parity] case={name} nE={mesh['nE']} kernel={kernel['function']} paths={paths}", flush=True)
    ref = oracle.par_system(mesh, kernel, conf, keep_neighbours=True)
    A, (ptr, nbr, cls), st = _gpu(mesh, kernel, conf, paths)
    # neighbour lists: identical sets per outer element
    assert np.array_equal(ptr, ref["nbr_ptr"]), "neighbour list lengths differ"
    for g, r in zip(sorted_lists(ptr, nbr), sorted_lists(ref["nbr_ptr"], ref["nbr"])):
        assert np.array_equal(g, r)
    # close (first-layer tensor Gauss) pairs are flagged identically
    if conf.get("closeElements") in ("fractional", "weakSingular"):
        gclose = sorted((i, b) for i in range(len(ptr) - 1) for b, c in
                        zip(nbr[ptr[i]:ptr[i + 1]], cls[ptr[i]:ptr[i + 1]]) if c == 2)
        rclose = sorted((i, b) for i in range(len(ptr) - 1) for b, c in
                        zip(ref["nbr"][ref["nbr_ptr"][i]:ref["nbr_ptr"][i + 1]],
                            ref["nbr_close"][ref["nbr_ptr"][i]:ref["nbr_ptr"][i + 1]]) if c)
        assert gclose == rclose
    assert st["pairs"] == ref["n_pairs"] and st["probes"] == ref["n_probes"]
    same, rel, rowrel = compare(A, ref["A"])
    print(f"[parity] case={name} pairs={int(st['pairs'])} nnz={A.nnz} pattern_equal={same} "
          f"max_rel={rel:.3e} max_rowrel={rowrel:.3e}", flush=True)
    assert same, "sparsity pattern differs"
    D = (A - ref["A"]).tocoo()
    Abs = oracle.abs_matrix(mesh, kernel, conf).tocsr()
    addends = np.asarray(Abs[D.row, D.col]).ravel()
    worst = float(np.max(np.abs(D.data) / np.maximum(addends, 1e-300))) if D.nnz else 0.0
    # plus 1e-15 of the largest entry: an entry made only of tiny basis-function products
    # (psi_a ~ 1e-5 near a far vertex of the inner element) carries an absolute rounding error
    # of the pair scale in the reference's own point-by-point evaluation as well
    floor = 1e-15 * abs(ref["A"]).max()
    bound = tol * addends + floor
    n_floor, worst_floorless = floor_usage(D.data, addends, tol, floor)
    print(f"[parity] case={name} max |gpu-oracle| / sum|addends| = {worst:.3e}; "
          f"max |gpu-oracle| / bound = {float(np.max(np.abs(D.data) / bound)) if D.nnz else 0.0:.3e}; "
          f"entries needing the floor: {n_floor} of {A.nnz}, their worst floorless ratio {worst_floorless:.3e}",
          flush=True)
    assert np.all(np.abs(D.data) <= bound), f"max rel {rel:.3e}"
    check_floor_usage(name, n_floor, worst_floorless, A.nnz)
    return A, ref


@pytest.mark.parametrize("name", ["C1", "C2s", "C3s", "C5s", "DG", "bary", "weak", "nocaps"])
def test_parity_cases(name):
    mesh, kernel, conf = case(name)
    A, ref = _check_case(name, mesh, kernel, conf)
    if kernel["outputdim"] == 2:
        _check_rigid(name, A, mesh, kernel, conf)


def _check_rigid(name, A, mesh, kernel, conf):
    """The GPU matrix annihilates the rigid motions in the reference's interleaved numbering."""
    res = rigid_residuals(A, oracle.abs_matrix(mesh, kernel, conf), mesh)
    print(f"[parity] case={name} rigid motions: {res}", flush=True)
    assert max(res.values()) < 1e-12, res


@pytest.mark.parametrize("name,art", [("hole", True), ("disjoint", True), ("disjoint", False), ("circle", True)])
def test_parity_tutorial_gmsh_meshes(name, art):
    """The reference tutorial's own meshes (examples/tutorial/data/*.msh) with the artificial
    vertex that joins every boundary line (128, 224 and 146 label-0 elements sharing one
    vertex, examples/tutorial/mesh.py:23-62): neighbour lists, pattern and values against the
    oracle.  circle runs the microelastic field kernel (outdim 2, 16/7 rules)."""
    mesh, kernel, conf = gmsh_case(name, set_art_vertex=art)
    A, ref = _check_case(f"{name}{'-av' if art else ''}", mesh, kernel, conf)
    if kernel["outputdim"] == 2:
        _check_rigid(name, A, mesh, kernel, conf)


@pytest.mark.parametrize("name", ["C1", "C3s", "C5s", "DG"])
def test_parity_domain_decomposition_weights(name):
    """nZeta > 0: every pair weighted by 1 / |Z_aT n Z_bT| (src/mathhelpers.cpp:19-38,
    src/Cassemble.cpp:648-656), on a random cover (every element in subdomain 0, about half in
    1, a fifth in 2), against the oracle: remote pairs (C1, DG), tensor-Gauss close pairs (C3s)
    and the field kernel (C5s)."""
    from test_oracle_zeta import random_cover, with_zeta
    mesh, kernel, conf = case(name)
    zmesh = with_zeta(mesh, random_cover(mesh["nE"], seed=11))
    A, ref = _check_case(f"{name}-zeta", zmesh, kernel, conf)
    plain = oracle.par_system(mesh, kernel, conf)["A"]
    assert abs(A - plain).max() > 0.05 * abs(plain).max()  # the weights act


def test_domain_decomposition_drop_in_and_load_vector(tmp_path):
    """The weights through par_assemble's file contract ("system" and "forcing")."""
    from test_oracle_zeta import random_cover, with_zeta
    mesh, kernel, conf = case("C2s")
    zmesh = with_zeta(mesh, random_cover(mesh["nE"], seed=3))
    conf2 = dict(conf)
    conf2["savePath"] = str(tmp_path / "A.arma")
    A = nlfem.stiffnessMatrix(zmesh, kernel, conf2)
    R = oracle.par_system(zmesh, kernel, conf)["A"]
    same, rel, rowrel = compare(A, R)
    print(f"[parity] zeta drop-in C2s pattern_equal={same} max_rowrel={rowrel:.3e}", flush=True)
    assert same and rowrel <= 1e-12
    f = nlfem.loadVector(zmesh, {"function": "linear"}, conf)
    fr = oracle.load_vector(zmesh, {"function": "linear"}, conf)
    f0 = oracle.load_vector(mesh, {"function": "linear"}, conf)
    assert np.allclose(f, fr, rtol=1e-13, atol=1e-15) and not np.allclose(f, f0)


@pytest.mark.parametrize("which", ["scalar30h", "field20h"])
def test_parity_large_horizons(which):
    """Horizons whose rows outgrow the LDS tables: 30 elements per horizon with the constant
    kernel (pattern element and node sets in global memory, cut pairs through the per-pair
    kernel), 20 with the field kernel (interior pairs through the per-pair kernel)."""
    from nlfem_amd.mesh import unit_square
    from nlfem_amd.config import make_configuration
    if which == "scalar30h":
        h = 1 / 32
        mesh = unit_square(64, 30 * h, jitter=0.2, seed=21, hi=2.0)
        kern = {"function": "constant", "horizon": 30 * h, "outputdim": 1}
        conf = make_configuration()
    else:
        h = 1 / 32
        mesh = unit_square(48, 20 * h, jitter=0.2, seed=22, hi=1.5, outdim=2)
        kern = {"function": "linearPrototypeMicroelasticField", "horizon": 20 * h, "outputdim": 2, "fractional_s": -0.5}
        conf = make_configuration(close="retriangulate", outer="16point", inner="7point")
    A, ref = _check_case(which, mesh, kern, conf)
    if kern["outputdim"] == 2:
        _check_rigid(which, A, mesh, kern, conf)


def test_parity_vertex_of_valence_140():
    """A fan of 140 triangles around one vertex: dual-graph rows of ~140 elements and a node
    row merged from 140 partial rows (more than one merge chunk)."""
    mesh, kernel, conf = fan_mesh(140)
    _check_case("fan140", mesh, kernel, conf)


def test_parity_label0_traversal():
    mesh, kernel, conf = label0_mesh()
    _check_case("label0", mesh, kernel, conf)


def test_parity_full_connected_component():
    from nlfem_amd.mesh import unit_square
    from nlfem_amd.config import make_configuration
    mesh = unit_square(6, 0.1)
    conf = make_configuration()
    conf["is_fullConnectedComponentSearch"] = 1
    _check_case("fullCC", mesh, {"function": "constant", "horizon": 0.1, "outputdim": 1}, conf)


def test_mixed_precision_tolerance():
    mesh, kernel, conf = case("C2s")
    conf = dict(conf)
    conf["precision"] = "mixed"
    ref = oracle.par_system(mesh, kernel, conf)
    A, _, _ = _gpu(mesh, kernel, conf)
    same, rel, rowrel = compare(A, ref["A"])
    fro = sp.linalg.norm(A - ref["A"]) / sp.linalg.norm(ref["A"])
    print(f"[parity] mixed C2s rel_fro={fro:.3e} max_rowrel={rowrel:.3e}", flush=True)
    assert same
    assert fro < 1e-5


def test_mixed_precision_fractional_C3s():
    """Config C3 in its stated precision: fractional kernel s = 0.4, close pairs by tensor Gauss
    of degree 5 in FP64, far pairs with FP32 quadrature sums (FP32 exp2/log2 of the kernel) and
    FP64 accumulation, against the FP64 oracle: identical pattern, relative Frobenius < 1e-5."""
    mesh, kernel, conf = case("C3s")
    from nlfem_amd.config import make_configuration
    conf = make_configuration(close="fractional", tensorGaussDegree=5)
    conf["precision"] = "mixed"
    ref = oracle.par_system(mesh, kernel, conf)
    A, _, st = _gpu(mesh, kernel, conf)
    same, rel, rowrel = compare(A, ref["A"])
    fro = sp.linalg.norm(A - ref["A"]) / sp.linalg.norm(ref["A"])
    print(f"[parity] mixed C3s degree 5 close={int(st['close'])} rel_fro={fro:.3e} max_rowrel={rowrel:.3e}",
          flush=True)
    assert st["close"] > 0
    assert same
    assert fro < 1e-5


def test_retriangulate_batch_bitexact():
    rng = np.random.default_rng(5)
    n = 4000
    TE = rng.uniform(0, 1, size=(n, 6))
    x = rng.uniform(-0.2, 1.2, size=(n, 2))
    sq = 0.3 ** 2
    # tangent and vertex-on-circle cases
    TE[:50] = [0, 0, 1, 0, 0, 1]
    x[:50] = np.stack([np.linspace(0.1, 0.9, 50), np.full(50, -0.3)], axis=1)
    TE[50:100] = [0, 0, 0.3, 0, 0, 0.5]
    x[50:100] = [0, 0]
    ntri, lists = nlfem.retriangulate_batch(x, TE, sq, caps=1)
    for i in range(n):
        r, l = oracle.retriangulate(x[i], TE[i], sq, caps=1)
        assert ntri[i] == r, i
        assert np.array_equal(lists[i, :6 * r], l[:6 * r]), i
    ntri0, lists0 = nlfem.retriangulate_batch(x, TE, sq, caps=0)
    for i in range(0, n, 7):
        r, l = oracle.retriangulate(x[i], TE[i], sq, caps=0)
        assert ntri0[i] == r and np.array_equal(lists0[i, :6 * r], l[:6 * r])


def test_geometry_helpers():
    E = np.array([0.1, 0.2, 0.9, 0.25, 0.3, 0.8])
    p = np.array([0.2, 0.3])
    x = nlfem.py_toPhys(E, p)
    assert np.array_equal(x, oracle.to_phys(E, p))
    assert np.array_equal(nlfem.py_toRef(E, x), oracle.to_ref(E, x))
    A = np.array([0.5, 2.0, 3.0, -1.0])
    b = np.array([1.0, 2.0])
    assert np.array_equal(nlfem.py_solve2x2(A, b), oracle.solve2x2(A, b))
    with pytest.raises(nlfem.NlfemError):
        nlfem.py_solve2x2(np.zeros(4), b)


def test_file_contract_matches_in_memory(tmp_path):
    mesh, kernel, conf = case("C1")
    A = nlfem.stiffnessMatrix(mesh, kernel, conf)
    conf2 = dict(conf)
    conf2["savePath"] = str(tmp_path / "A.arma")
    B = nlfem.stiffnessMatrix(mesh, kernel, conf2)
    with open(conf2["savePath"], "rb") as f:
        assert f.readline() == b"ARMA_SPM_BIN_FN008\n"
    # same pattern; values agree to rounding (FP64 atomics sum in a run-dependent order)
    assert np.array_equal(A.indptr, B.indptr) and np.array_equal(A.indices, B.indices)
    assert abs(A - B).max() <= 1e-14 * abs(A).max()


def test_load_vector_and_mass():
    mesh, kernel, conf = case("C2s")
    f = nlfem.loadVector(mesh, {"function": "linear"}, conf)
    fr = oracle.load_vector(mesh, {"function": "linear"}, conf)
    assert np.allclose(f, fr, rtol=1e-13, atol=1e-15)
    u = np.random.default_rng(0).normal(size=mesh["nV"])
    v = nlfem.evaluateMass(mesh, u, conf["quadrature"]["outer"]["points"], conf["quadrature"]["outer"]["weights"],
                           J=mesh["nE"])
    vr = oracle.evaluate_mass(mesh, u, conf["quadrature"]["outer"]["points"], conf["quadrature"]["outer"]["weights"])
    assert np.allclose(v, vr, rtol=1e-12, atol=1e-15)


def test_invalid_configuration_raises():
    mesh, kernel, conf = case("C1")
    with pytest.raises(nlfem.NlfemError):
        nlfem.stiffnessMatrix(mesh, {"function": "nope", "horizon": 0.1, "outputdim": 1}, conf)
    bad = dict(conf)
    bad["approxBalls"] = {"method": "fractional"}
    with pytest.raises(nlfem.NlfemError):
        nlfem.stiffnessMatrix(mesh, kernel, bad)


def test_parity_shuffled_element_numbering():
    """Elements in random order: the dual graph's index bandwidth is large, so the BFS uses
    the hashed visited set instead of the index-window bitmap."""
    from nlfem_amd.mesh import unit_square, mesh_dict
    from nlfem_amd.config import make_configuration
    m = unit_square(48, 0.1, jitter=0.1, seed=9)
    perm = np.random.default_rng(3).permutation(m["nE"])
    ms = mesh_dict(m["elements"][perm], m["elementLabels"][perm], m["vertices"], m["vertexLabels"])
    _check_case("shuffled", ms, {"function": "constant", "horizon": 0.1, "outputdim": 1}, make_configuration())


@pytest.mark.gpu
def test_parity_general_paths():
    """The hashed BFS visited sets, hashed pattern element sets, counted CSR transposed rows and
    cut entries of the transposed rows placed one at a time (also on baryCenter pairs, whose
    entries carry bit 31), forced on a locally numbered mesh, give the same lists, pattern and
    values."""
    from nlfem_amd.mesh import unit_square
    from nlfem_amd.config import make_configuration
    m = unit_square(40, 0.1, jitter=0.15, seed=5)
    flags = nlfem.Assembler.PATH_BFS_HASH | nlfem.Assembler.PATH_PATTERN_HASH | nlfem.Assembler.PATH_TRANSPOSE_CSR
    _check_case("general-paths", m, {"function": "constant", "horizon": 0.1, "outputdim": 1}, make_configuration(),
                paths=flags)
    _check_case("csr-transpose", m, {"function": "parabola", "horizon": 0.1, "outputdim": 1}, make_configuration(),
                paths=nlfem.Assembler.PATH_TRANSPOSE_CSR)
    _check_case("bfs-wave", m, {"function": "constant", "horizon": 0.1, "outputdim": 1}, make_configuration(),
                paths=nlfem.Assembler.PATH_BFS_WAVE)
    _check_case("transpose-unstaged", m, {"function": "constant", "horizon": 0.1, "outputdim": 1},
                make_configuration(), paths=nlfem.Assembler.PATH_TRANSPOSE_UNSTAGED)
    mb, kb, cb = case("bary")  # baryCenter pairs: cut entries with bit 31
    _check_case("bary-transpose-unstaged", mb, kb, cb, paths=nlfem.Assembler.PATH_TRANSPOSE_UNSTAGED)
    _check_case("bary-transpose-csr", mb, kb, cb, paths=nlfem.Assembler.PATH_TRANSPOSE_CSR)


@pytest.mark.parametrize("name", ["C5s", "DG", "C2s"])
def test_cut_items_per_chunk_and_autotune(name):
    """The general cut kernel with 64, 128 and 256 items per chunk, the per-lane cut kernel
    (items 0), and the plan after the autotune meet the FP64 bar against the oracle (the choice
    changes summation order only)."""
    mesh, kernel, conf = case(name)
    ref = oracle.par_system(mesh, kernel, conf)
    Abs = oracle.abs_matrix(mesh, kernel, conf).tocsr()
    floor = 1e-15 * abs(ref["A"]).max()

    def check(A, label):
        same, rel, _ = compare(A, ref["A"])
        D = (A - ref["A"]).tocoo()
        addends = np.asarray(Abs[D.row, D.col]).ravel()
        bound = FP64_TOL * addends + floor
        n_floor, worst_floorless = floor_usage(D.data, addends, FP64_TOL, floor)
        print(f"[parity] case={name} {label} pattern_equal={same} max_rel={rel:.3e} floor entries={n_floor} "
              f"(worst floorless {worst_floorless:.3e})", flush=True)
        assert same and np.all(np.abs(D.data) <= bound), label
        check_floor_usage(f"{name} {label}", n_floor, worst_floorless, A.nnz)

    for items in (64, 128, 256, 0):
        asm = nlfem.Assembler(mesh, kernel, conf).set_cut_items(items)
        try:
            check(asm.assemble().csr(), f"items={items}")
        finally:
            asm.close()
    asm = nlfem.Assembler(mesh, kernel, conf)
    try:
        chosen, ms = asm.autotune()
        print(f"[parity] case={name} autotune chose {chosen}, cut ms {ms}", flush=True)
        assert chosen in (0, 64, 128, 256) and ms[128] > 0 and ms[0] > 0 and ms[chosen] > 0
        check(asm.csr(), "after autotune")
        check(asm.assemble().csr(), f"tuned items={chosen}")
    finally:
        asm.close()
    asm = nlfem.Assembler(mesh, kernel, conf)
    try:
        with pytest.raises(nlfem.NlfemError):
            asm.set_cut_items(100)
    finally:
        asm.close()


@pytest.mark.gpu
def test_parity_odd_element_count():
    """An odd number of elements: the BFS with two outer elements per wavefront has an idle
    second half in its last wavefront.  The square mesh minus one interior element (a hole)
    gives 199 elements; lists, pattern and values match the oracle, on the bitmap BFS and on
    the one-element-per-wavefront kernel."""
    from nlfem_amd.mesh import mesh_dict, unit_square
    from nlfem_amd.config import make_configuration
    m = unit_square(10, 0.2, jitter=0.1, seed=3)
    keep = np.ones(m["nE"], dtype=bool)
    keep[105] = False
    odd = mesh_dict(m["elements"][keep], m["elementLabels"][keep], m["vertices"], m["vertexLabels"])
    assert odd["nE"] % 2 == 1
    kern = {"function": "constant", "horizon": 0.2, "outputdim": 1}
    _check_case("odd", odd, kern, make_configuration())
    _check_case("odd-wave", odd, kern, make_configuration(), paths=nlfem.Assembler.PATH_BFS_WAVE)


def _rule12():
    """A 12-point outer rule: the 3-point degree-2 rule on each of the four midpoint
    subtriangles of the reference triangle (weights 1/24, sum 1/2; chk_BasisFunction holds)."""
    base = np.array([[1 / 6, 1 / 6], [2 / 3, 1 / 6], [1 / 6, 2 / 3]])
    subs = [np.array([[0, 0], [0.5, 0], [0, 0.5]]), np.array([[0.5, 0], [1, 0], [0.5, 0.5]]),
            np.array([[0, 0.5], [0.5, 0.5], [0, 1]]), np.array([[0.5, 0.5], [0, 0.5], [0.5, 0]])]
    pts = []
    for T in subs:
        for p in base:
            pts.append(T[0] + p[0] * (T[1] - T[0]) + p[1] * (T[2] - T[0]))
    return np.array(pts), np.full(12, 1 / 24)


@pytest.mark.parametrize("kern", ["parabola", "linearPrototypeMicroelasticField"])
def test_parity_outer_rule_12_points(kern):
    """Outer rules of 9..15 points take the per-lane cut kernel with groups of nPx lanes whose
    pair sums go through LDS atomics (the 8- and 16-lane groups reduce in registers): a
    12-point outer rule on a scalar and on the field kernel against the oracle."""
    od = 2 if kern.endswith("Field") else 1
    mesh = unit_square(12, 0.15, jitter=0.2, seed=5, outdim=od)
    kernel = {"function": kern, "horizon": 0.15, "outputdim": od, "fractional_s": -0.5}
    conf = make_configuration(close="retriangulate")
    Px, dx = _rule12()
    assert abs(dx.sum() - 0.5) < 1e-15
    conf["quadrature"]["outer"] = {"points": Px, "weights": dx}
    A, ref = _check_case(f"rule12-{kern}", mesh, kernel, conf)
    if od == 2:
        _check_rigid(f"rule12-{kern}", A, mesh, kernel, conf)


@pytest.mark.gpu
@pytest.mark.parametrize("n,delta,kern,jitter,ckw", [
    (256, 0.05, {"function": "constant", "horizon": 0.05, "outputdim": 1}, 0.0, {}),
    (128, 0.1, {"function": "parabola", "horizon": 0.1, "outputdim": 1}, 0.2, {}),
    (64, 0.2, {"function": "fractional", "horizon": 0.2, "outputdim": 1, "fractional_s": 0.4}, 0.0,
     {"close": "fractional", "tensorGaussDegree": 5}),
], ids=["h256_constant", "h128_parabola_jitter", "h64_fractional"])
def test_full_size_invariants(n, delta, kern, jitter, ckw):
    """At the full sizes of the single-mesh workloads (the oracle takes 4 to 40 s on them; the
    bench's --vs-oracle runs compare them entry by entry): A 1 = 0 on every assembled row and a
    symmetric Omega x Omega block to 1e-13 of max|A| (an entry without its mirror counts with its
    whole value: exact zeros are dropped, so the pattern itself may differ by cancellation
    residues), the oracle's own bound on the small cases (test_oracle.test_matrix_invariants).
    Measured: row sums <= 1.1e-15 and asymmetry <= 1.1e-16 of max|A|; a dropped term or a wrong
    index breaks them by orders of magnitude more."""
    m = unit_square(n, delta, jitter=jitter, seed=2024)
    print(f"[full size] n={n} delta={delta} kernel={kern['function']} nE={m['nE']}", flush=True)
    a = nlfem.Assembler(m, kern, make_configuration(**ckw))
    try:
        A = a.assemble().csr()
    finally:
        a.close()
    scale = abs(A).max()
    rows = np.where(m["nodeLabels"] > 0)[0]
    rowsum = np.abs(np.asarray(A[rows].sum(axis=1))).max()
    AO = A[rows][:, rows].tocsr()
    D = (AO - AO.T).tocsr()
    asym = abs(D).max() if D.nnz else 0.0
    print(f"[full size] nnz={A.nnz} max|A|={scale:.3e} row sums {rowsum / scale:.2e} "
          f"asymmetry {asym / scale:.2e}", flush=True)
    assert rowsum <= 1e-13 * scale
    assert asym <= 1e-13 * scale


@pytest.mark.gpu
def test_full_size_field_kernel_rigid_motions():
    """The microelastic field kernel (outdim 2, 16/7 rules) on a jittered h=1/128 mesh with
    delta = 0.1 (12.8 h, the batch workload's largest horizon per element): translations and the
    rotation (interleaved degrees of freedom, src/Cassemble.cpp:736-766) lie in the null space
    of every assembled row, relative to (|A| |u|)_r, to 1e-9 (identical pairs sum addends ~1e6
    times the entries; the oracle's own bound on the small field case is 1e-8 of max|A|).
    Measured: 1.7e-11 for the translations, 1.1e-11 for the rotation."""
    m = unit_square(128, 0.1, jitter=0.2, seed=2025, outdim=2)
    kern = {"function": "linearPrototypeMicroelasticField", "horizon": 0.1, "outputdim": 2, "fractional_s": -0.5}
    conf = make_configuration(close="retriangulate", outer="16point", inner="7point")
    print(f"[full size field] nE={m['nE']}", flush=True)
    a = nlfem.Assembler(m, kern, conf)
    try:
        A = a.assemble().csr()
    finally:
        a.close()
    res = rigid_residuals(A, abs(A), m)
    print(f"[full size field] nnz={A.nnz} rigid-motion residuals {res}", flush=True)
    assert max(res.values()) <= 1e-9
