"""ctypes wrapper of the CPU oracle (oracle/nlfem_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker.  The product (nlfem_amd) never imports this module.
"""
import ctypes as C
import os

import numpy as np
import scipy.sparse as sp

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = C.POINTER(C.c_double)
_lp = C.POINTER(C.c_long)
_ip = C.POINTER(C.c_int)


def lib_path():
    return os.path.join(_HERE, "libnlfem_oracle.so")


def load():
    global _LIB
    if _LIB is not None:
        return _LIB
    path = lib_path()
    if not os.path.exists(path):
        raise RuntimeError(f"oracle library {path} is not built; run `make -C oracle` or __graft_entry__.build()")
    lib = C.CDLL(path)
    lib.oracle_par_system.restype = C.c_void_p
    lib.oracle_par_system.argtypes = [C.c_int, _lp, _lp, _dp, _lp, C.c_int, C.c_int, _dp, C.c_int, _dp, _dp, C.c_int,
                                      _dp, C.c_double, C.c_int, C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, C.c_int,
                                      _dp, C.c_int, _dp, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                      _lp, _lp, C.c_long, C.c_int]
    for name in ("oracle_result_nnz", "oracle_result_npairs", "oracle_result_nprobes", "oracle_result_nclose",
                 "oracle_result_nbr_total"):
        getattr(lib, name).restype = C.c_long
        getattr(lib, name).argtypes = [C.c_void_p]
    lib.oracle_result_seconds.restype = C.c_double
    lib.oracle_result_seconds.argtypes = [C.c_void_p]
    lib.oracle_result_csr.argtypes = [C.c_void_p, _lp, _lp, _dp]
    lib.oracle_result_neighbours.argtypes = [C.c_void_p, _lp, _ip, C.POINTER(C.c_ubyte)]
    lib.oracle_result_free.argtypes = [C.c_void_p]
    lib.oracle_set_abs_mode.argtypes = [C.c_int]
    lib.oracle_set_plant.argtypes = [C.c_int]
    lib.oracle_last_error.restype = C.c_char_p
    lib.oracle_dual_graph.restype = C.c_long
    lib.oracle_dual_graph.argtypes = [_lp, C.c_int, C.c_int, _ip, _ip]
    lib.oracle_method_retriangulate.argtypes = [_dp, _dp, C.c_double, _dp, C.c_int]
    lib.oracle_toRef.argtypes = [_dp, _dp, _dp]
    lib.oracle_toPhys.argtypes = [_dp, _dp, _dp]
    lib.oracle_solve2x2.argtypes = [_dp, _dp, _dp]
    lib.oracle_kernel.argtypes = [C.c_char_p, _dp, _dp, C.c_double, C.c_double, _dp]
    lib.oracle_integrate_pair.argtypes = [C.c_int, _lp, _dp, C.c_int, C.c_int, C.c_int, _dp, C.c_int, _dp, _dp,
                                          C.c_int, _dp, C.c_double, C.c_char_p, C.c_int, C.c_int, _dp, C.c_int, _dp,
                                          C.c_double, _dp, _dp, _dp, _dp]
    lib.oracle_quadrule.argtypes = [_dp, C.c_int, C.c_double, C.c_int, _dp, _dp, _dp, _dp]
    lib.oracle_par_forcing.argtypes = [C.c_int, _lp, _lp, _dp, _lp, C.c_int, _dp, C.c_int, _dp, C.c_int, C.c_char_p,
                                       C.c_int, _dp, _lp, _lp, C.c_long]
    lib.oracle_par_evaluateMass.argtypes = [_dp, _dp, _lp, _lp, _dp, _lp, C.c_int, C.c_int, _dp, _dp, C.c_int, C.c_int]
    _LIB = lib
    return lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _l(a):
    return a.ctypes.data_as(_lp)


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel())


def _i64(a):
    return np.ascontiguousarray(np.asarray(a).astype(np.int64).ravel())


def _zeta(mesh):
    """Domain-decomposition indicator of the mesh dictionary as the reference's wrapper reads it
    (cython/nlfem.pyx:205-214): mesh["nZeta"] and the CSR rows of mesh["ZetaIndicator"]."""
    n = int(mesh.get("nZeta", 0))
    if n <= 0:
        z = np.zeros(1, dtype=np.int64)
        return z, z, 0
    Z = mesh["ZetaIndicator"]
    return _i64(Z.indices), _i64(Z.indptr), n


def par_system(mesh, kernel, conf, nthreads=0, keep_neighbours=False, outer=None, stride=1):
    """Assemble with the oracle.  Returns a dict with ``A`` (scipy CSR, K x K) and counters.

    ``outer=(begin, end)`` restricts the outer-element loop to a range and ``stride`` to every
    stride-th outer element (samples for timing)."""
    from nlfem_amd.config import parse
    a = parse(mesh, kernel, conf)
    zi, zp, nz = _zeta(mesh)
    lib = load()
    Pg = a.Pg if a.Pg.size else np.zeros(1)
    dg = a.dg if a.dg.size else np.zeros(1)
    h = lib.oracle_par_system(a.K, _l(a.elements), _l(a.elementLabels), _d(a.vertices), _l(a.vertexLabels), a.nE,
                              a.nV, _d(a.Px), a.Px.size // 2, _d(a.dx), _d(a.Py), a.Py.size // 2, _d(a.dy),
                              a.sqdelta, a.is_DG, a.kernel.encode(), a.method_remote.encode(),
                              a.method_close.encode(), a.is_PlacePointOnCap, a.outdim, _d(Pg), a.degree, _d(dg),
                              a.fractional_s, a.is_fullConnectedComponentSearch, int(nthreads), int(keep_neighbours),
                              int(outer[0]) if outer else 0, int(outer[1]) if outer else 0, _l(zi), _l(zp), nz,
                              int(stride))
    if not h:
        raise ValueError("oracle: " + lib.oracle_last_error().decode())
    try:
        nnz = lib.oracle_result_nnz(h)
        rowptr = np.zeros(a.K + 1, dtype=np.int64)
        col = np.zeros(max(nnz, 1), dtype=np.int64)
        val = np.zeros(max(nnz, 1), dtype=np.float64)
        lib.oracle_result_csr(h, _l(rowptr), _l(col), _d(val))
        A = sp.csr_matrix((val[:nnz], col[:nnz], rowptr), shape=(a.K, a.K))
        out = {"A": A, "n_pairs": lib.oracle_result_npairs(h), "n_probes": lib.oracle_result_nprobes(h),
               "n_close": lib.oracle_result_nclose(h), "seconds": lib.oracle_result_seconds(h)}
        if keep_neighbours:
            tot = lib.oracle_result_nbr_total(h)
            ptr = np.zeros(a.nE + 1, dtype=np.int64)
            nbr = np.zeros(max(tot, 1), dtype=np.int32)
            cl = np.zeros(max(tot, 1), dtype=np.uint8)
            lib.oracle_result_neighbours(h, _l(ptr), nbr.ctypes.data_as(_ip), cl.ctypes.data_as(C.POINTER(C.c_ubyte)))
            out["nbr_ptr"], out["nbr"], out["nbr_close"] = ptr, nbr[:tot], cl[:tot]
        return out
    finally:
        lib.oracle_result_free(h)


def abs_matrix(mesh, kernel, conf):
    """Sum of |addend| per entry (same pattern as par_system): a bound for rounding effects."""
    lib = load()
    lib.oracle_set_abs_mode(1)
    try:
        return par_system(mesh, kernel, conf)["A"]
    finally:
        lib.oracle_set_abs_mode(0)


class planted:
    """Context manager that plants a known mistake in the oracle (tests that must show they
    catch it): "blocked_layout" numbers outdim-2 degrees of freedom comp * nV + node instead of
    2 node + comp; "field_diagonal" exchanges k00 and k11 of the field kernel."""
    MODES = {"blocked_layout": 1, "field_diagonal": 2}

    def __init__(self, mode):
        self.mode = self.MODES[mode]

    def __enter__(self):
        load().oracle_set_plant(self.mode)
        return self

    def __exit__(self, *exc):
        load().oracle_set_plant(0)
        return False


def dual_graph(elements, nV):
    lib = load()
    el = _i64(elements)
    nE = el.size // 3
    xadj = np.zeros(nE + 1, dtype=np.int32)
    n = lib.oracle_dual_graph(_l(el), nE, nV, xadj.ctypes.data_as(_ip), None)
    adj = np.zeros(max(n, 1), dtype=np.int32)
    lib.oracle_dual_graph(_l(el), nE, nV, xadj.ctypes.data_as(_ip), adj.ctypes.data_as(_ip))
    return xadj, adj[:n]


def retriangulate(x, TE, sqdelta, caps=1):
    lib = load()
    x, TE = _f64(x), _f64(TE)
    out = np.zeros(9 * 3 * 2)
    n = lib.oracle_method_retriangulate(_d(x), _d(TE), float(sqdelta), _d(out), int(caps))
    return n, out


def to_ref(E, x):
    lib = load()
    E, x = _f64(E), _f64(x)
    p = np.zeros(2)
    if lib.oracle_toRef(_d(E), _d(x), _d(p)):
        raise ValueError("singular element")
    return p


def to_phys(E, p):
    lib = load()
    E, p = _f64(E), _f64(p)
    x = np.zeros(2)
    lib.oracle_toPhys(_d(E), _d(p), _d(x))
    return x


def solve2x2(A, b):
    lib = load()
    A, b = _f64(A), _f64(b)
    x = np.zeros(2)
    if lib.oracle_solve2x2(_d(A), _d(b), _d(x)):
        raise ValueError("singular matrix")
    return x


def kernel_value(name, x, y, sqdelta, fractional_s=-1.0):
    lib = load()
    x, y = _f64(x), _f64(y)
    out = np.zeros(4)
    if lib.oracle_kernel(name.encode(), _d(x), _d(y), float(sqdelta), float(fractional_s), _d(out)):
        raise ValueError(name)
    return out


METHOD_IDS = {"retriangulate": 0, "baryCenter": 1, "fractional": 2, "weakSingular": 3}


def integrate_pair(method, elements, vertices, aTdx, bTdx, Px, dx, Py, dy, sqdelta, kernel, caps=1, outdim=1,
                   degree=0, fractional_s=-1.0):
    """Four local blocks (termLocal, termNonloc, termLocalPrime, termNonlocPrime) of one pair."""
    from nlfem_amd.quadrature import tensorgauss
    lib = load()
    el, vt = _i64(elements), _f64(vertices)
    Px, dx, Py, dy = _f64(Px), _f64(dx), _f64(Py), _f64(dy)
    if degree:
        tg = tensorgauss(degree)
        Pg, dg = _f64(tg.points), _f64(tg.weights)
    else:
        Pg, dg = np.zeros(1), np.zeros(1)
    nb = 9 * outdim * outdim
    blocks = [np.zeros(nb) for _ in range(4)]
    r = lib.oracle_integrate_pair(METHOD_IDS[method], _l(el), _d(vt), vt.size // 2, int(aTdx), int(bTdx), _d(Px),
                                  Px.size // 2, _d(dx), _d(Py), Py.size // 2, _d(dy), float(sqdelta), kernel.encode(),
                                  int(caps), int(outdim), _d(Pg), int(degree), _d(dg), float(fractional_s),
                                  *[_d(b) for b in blocks])
    n = 3 * outdim
    return r, [b.reshape(n, n) for b in blocks]


def quadrule_tables(degree, s, case):
    from nlfem_amd.quadrature import tensorgauss
    lib = load()
    tg = tensorgauss(degree)
    Pg = _f64(tg.points)
    nT = (2, 5, 6)[case]
    nPg = degree ** 4
    al = np.zeros(nT * nPg * 4)
    ac = np.zeros(nT * nPg * 4)
    df = np.zeros(nT * nPg)
    dw = np.zeros(nT * nPg)
    lib.oracle_quadrule(_d(Pg), degree, float(s), case, _d(al), _d(ac), _d(df), _d(dw))
    return (al.reshape(nT, nPg, 4), ac.reshape(nT, nPg, 4), df.reshape(nT, nPg), dw.reshape(nT, nPg),
            _f64(tg.weights))


def load_vector(mesh, load_fn, conf):
    from nlfem_amd.config import parse
    lib = load()
    kern = {"function": "constant", "horizon": 0.1, "outputdim": mesh.get("outdim", 1)}
    a = parse(mesh, kern, conf)
    fd = np.zeros(a.K)
    zi, zp, nz = _zeta(mesh)
    r = lib.oracle_par_forcing(a.K, _l(a.elements), _l(a.elementLabels), _d(a.vertices), _l(a.vertexLabels), a.nE,
                               _d(a.Px), a.Px.size // 2, _d(a.dx), a.is_DG, load_fn["function"].encode(), a.outdim,
                               _d(fd), _l(zi), _l(zp), nz)
    if r:
        raise ValueError("unknown forcing " + load_fn["function"])
    return fd


def evaluate_mass(mesh, ud, Px, dx, J=None):
    lib = load()
    el, lab, vt, vlab = _i64(mesh["elements"]), _i64(mesh["elementLabels"]), _f64(mesh["vertices"]), \
        _i64(mesh["vertexLabels"])
    Px, dx = _f64(Px), _f64(dx)
    ud = _f64(ud)
    vd = np.zeros(mesh["K"])
    J = mesh["nE"] if J is None else J
    lib.oracle_par_evaluateMass(_d(vd), _d(ud), _l(el), _l(lab), _d(vt), _l(vlab), int(J), Px.size // 2, _d(Px),
                                _d(dx), int(mesh.get("outdim", 1)), int(mesh.get("is_DiscontinuousGalerkin", 0)))
    return vd
