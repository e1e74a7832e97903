"""Python API of the reference (cython/nlfem.pyx), backed by libnlfem_amd.so on the GPU.

Same function names, dictionary keys, defaults and error behaviour as the reference module:

* :func:`stiffnessMatrix` (cython/nlfem.pyx:128-274) -> scipy CSR matrix A (K x K)
* :func:`loadVector` (:276-389), :func:`evaluateMass` (:638-676)
* :func:`read_arma_spMat`, :func:`read_arma_mat` (:28-101), :class:`tensorgauss` (:108-126)
* :func:`meshFromArrays`, :func:`get_vertexLabel`, :func:`setSolutionLabels`,
  :func:`stiffnessMatrix_fromArray` (:392-481)
* :func:`py_retriangulate`, :func:`py_toRef`, :func:`py_toPhys`, :func:`py_solve2x2` (:711-771)

With ``configuration["savePath"]`` set, :func:`stiffnessMatrix` goes through the reference's
file contract (``par_assemble`` writes ARMA_SPM_BIN_FN008, the wrapper reads it back);
otherwise the CSR is copied straight from device memory.  Additional configuration keys:
``"precision"``: ``"fp64"`` (default) or ``"mixed"`` (FP32 quadrature sums, FP64 accumulation
and geometry).
"""
import ctypes as C
import datetime
import os
import time

import numpy as np
import scipy.sparse as sp

from ._lib import lib, check, dp, lp, NlfemError, _fp
from .config import parse
from .quadrature import tensorgauss  # noqa: F401  (re-exported, cython/nlfem.pyx:108-126)
from .mesh import mesh_dict


def timestamp():
    return datetime.datetime.now().strftime("%m%d_%H-%M-%S")


def read_arma_mat(path, is_verbose=False):
    """Armadillo dense binary (ARMA_MAT_BIN_FN008) -> ndarray (cython/nlfem.pyx:28-58)."""
    with open(path, "rb") as f:
        if f.readline() != b"ARMA_MAT_BIN_FN008\n":
            raise ValueError("in read_arma_mat(), input file is of wrong format.")
        n_rows, n_cols = tuple(int(x) for x in f.readline().decode("utf-8").split())
        data = f.read()
    return np.array(np.frombuffer(data[:8 * n_rows * n_cols])).reshape((n_rows, n_cols), order="F")


def read_arma_spMat(path, is_verbose=False):
    """Armadillo sparse binary (ARMA_SPM_BIN_FN008, CSC of A^T) -> scipy CSR of A (cython/nlfem.pyx:60-101)."""
    with open(path, "rb") as f:
        if f.readline() != b"ARMA_SPM_BIN_FN008\n":
            raise ValueError("in read_arma_spMat(), input file is of wrong format.")
        n_rows, n_cols, nnz = tuple(int(x) for x in f.readline().decode("utf-8").split())
        data = f.read()
    values = np.frombuffer(data[:8 * nnz])
    pointers = np.frombuffer(data[8 * nnz:], dtype=np.uint64)
    row_index = pointers[:nnz]
    col_pointer = pointers[nnz:]
    A = sp.csc_matrix((values, row_index, col_pointer), shape=(n_rows, n_cols)).transpose()
    return A.tocsr()


def remove_arma_tmp(path):
    os.remove(path)


_PRECISION = {"fp64": 0, "mixed": 1}


def _strs(*s):
    return [x.encode("utf-8") for x in s]


def _zeta(mesh):
    """Domain-decomposition indicator as the reference's wrapper reads it
    (cython/nlfem.pyx:205-214): mesh["nZeta"] (0 or absent: no weights) and the CSR rows of
    mesh["ZetaIndicator"] (elements x subdomains)."""
    n = int(mesh.get("nZeta", 0)) if isinstance(mesh, dict) else 0
    if n <= 0:
        z = np.zeros(1, dtype=np.int64)
        return z, z, 0
    Z = mesh["ZetaIndicator"]
    zi = np.ascontiguousarray(np.asarray(Z.indices), dtype=np.int64)
    zp = np.ascontiguousarray(np.asarray(Z.indptr), dtype=np.int64)
    if zp.size != mesh["elements"].shape[0] + 1:
        raise ValueError("ZetaIndicator must have one row per element")
    return (zi if zi.size else np.zeros(1, dtype=np.int64)), zp, n


class Assembler:
    """Device-resident assembly plan (nlfem_plan_* in include/nlfem_amd.h).

    The mesh and configuration are uploaded once; :meth:`assemble` runs the whole pipeline on
    the GPU; the CSR stays in HBM until :meth:`csr` copies it out.
    """

    def __init__(self, mesh, kernel, configuration):
        a = parse(mesh, kernel, configuration)
        if configuration.get("precision", "fp64") not in _PRECISION:
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.args = a
        L = lib()
        h = C.c_void_p()
        Pg = a.Pg if a.Pg.size else np.zeros(1)
        dg = a.dg if a.dg.size else np.zeros(1)
        ks, mr, mc = _strs(a.kernel, a.method_remote, a.method_close)
        check(L.nlfem_plan_create(C.byref(h), a.K, lp(a.elements), lp(a.elementLabels), dp(a.vertices),
                                  lp(a.vertexLabels), a.nE, a.nV, dp(a.Px), a.Px.size // 2, dp(a.dx), dp(a.Py),
                                  a.Py.size // 2, dp(a.dy), a.sqdelta, a.is_DG, ks, mr, mc, a.is_PlacePointOnCap,
                                  a.outdim, dp(Pg), a.degree, dp(dg), a.fractional_s,
                                  a.is_fullConnectedComponentSearch,
                                  _PRECISION[configuration.get("precision", "fp64")]))
        self._h = h
        zi, zp, nz = _zeta(mesh)
        if nz > 0:
            check(L.nlfem_plan_set_zeta(h, lp(zi), lp(zp), nz))
        sur = configuration.get("surrogate")
        if sur:
            self.set_surrogate(**(sur if isinstance(sur, dict) else {}))

    PATH_BFS_HASH, PATH_PATTERN_HASH, PATH_TRANSPOSE_CSR, PATH_BFS_WAVE = 1, 2, 4, 8
    PATH_TRANSPOSE_UNSTAGED = 16

    def force_paths(self, flags):
        """Use the general data structures (hashed BFS visited sets, hashed pattern element
        sets, counted CSR transposed rows) that are otherwise taken only for poorly numbered
        meshes or after a fixed-size layout overflowed; PATH_BFS_WAVE runs the bitmap BFS with
        one outer element per wavefront instead of two; PATH_TRANSPOSE_UNSTAGED places the cut
        entries of the transposed rows one at a time (the path of blocks whose cut entries do
        not fit the LDS staging array)."""
        check(lib().nlfem_plan_force_paths(self._h, int(flags)))
        return self

    def set_cut_items(self, items):
        """Items (pair x outer point) per chunk of the general cut kernel: 64, 128 or 256, or
        0 for the per-lane kernel (one item per lane, no chunk barriers)."""
        check(lib().nlfem_plan_set_cut_items(self._h, int(items)))
        return self

    def autotune(self):
        """Time the general cut kernel with 64, 128 and 256 items per chunk and the per-lane
        kernel (0), one assembly each where the kernel runs and fits, and keep the fastest.
        Returns the kept value and the cut-phase milliseconds per candidate (0.0 where not
        tried)."""
        chosen = C.c_int(0)
        ms = np.zeros(4)
        check(lib().nlfem_plan_autotune(self._h, C.byref(chosen), ms.ctypes.data_as(C.POINTER(C.c_double))))
        return chosen.value, {64: float(ms[0]), 128: float(ms[1]), 256: float(ms[2]), 0: float(ms[3])}

    def set_surrogate(self, model="default", tol=None, smin=None, smax=None, enabled=True):
        """Route cut pairs through the surrogate networks of `model` (a path written by
        nlfem_amd.surrogate.save, or "default"); pairs outside the networks' domain or with an
        ensemble spread above `tol` stay with the exact retriangulation."""
        from . import surrogate
        if not enabled:
            check(lib().nlfem_plan_set_surrogate(self._h, 0, None, None, 0.0, 0.0, 1.0))
            return self
        params, norm, meta = surrogate.load(model, kernel=self.args.kernel)
        tol = float(meta.get("tol", 1e-3) if tol is None else tol)
        smin = float(meta.get("smin", 0.0) if smin is None else smin)
        smax = float(meta.get("smax", 1.0) if smax is None else smax)
        p = np.ascontiguousarray(params, dtype=np.float32)
        nm = np.ascontiguousarray(norm, dtype=np.float32)
        check(lib().nlfem_plan_set_surrogate(self._h, p.shape[0], p.ctypes.data_as(_fp), nm.ctypes.data_as(_fp),
                                             tol, smin, smax))
        return self

    def surrogate_samples(self, maxn, stride=1, salt=0, smin=0.0, smax=1.0):
        """Training samples (features [n,12], targets [n,9], or [n,18] for the microelastic field
        kernel) from the exact FP64 cut-pair sums of the assembled plan, and the number of
        qualifying items."""
        F = np.zeros((max(maxn, 1), 12), dtype=np.float32)
        T = np.zeros((max(maxn, 1), 18 if self.args.outdim == 2 else 9), dtype=np.float32)
        cnt = C.c_longlong()
        check(lib().nlfem_plan_surrogate_samples(self._h, int(maxn), int(stride), int(salt) & 0xFFFFFFFF,
                                                 float(smin), float(smax), F.ctypes.data_as(_fp),
                                                 T.ctypes.data_as(_fp), C.byref(cnt)))
        m = min(cnt.value, maxn)
        return F[:m], T[:m], cnt.value

    def assemble(self):
        check(lib().nlfem_plan_assemble(self._h))
        return self

    def csr(self):
        L = lib()
        K, nnz = C.c_longlong(), C.c_longlong()
        check(L.nlfem_plan_csr_info(self._h, C.byref(K), C.byref(nnz)))
        rowptr = np.zeros(K.value + 1, dtype=np.int64)
        col = np.zeros(max(nnz.value, 1), dtype=np.int32)
        val = np.zeros(max(nnz.value, 1), dtype=np.float64)
        check(L.nlfem_plan_csr_copy(self._h, rowptr.ctypes.data_as(C.POINTER(C.c_longlong)),
                                    col.ctypes.data_as(C.POINTER(C.c_int)), dp(val)))
        n = nnz.value
        return sp.csr_matrix((val[:n], col[:n], rowptr), shape=(K.value, K.value))

    def neighbours(self):
        """(ptr[nE+1], nbr[total], cls[total]) -- per outer element, unordered."""
        L = lib()
        tot = C.c_longlong()
        check(L.nlfem_plan_neighbours_total(self._h, C.byref(tot)))
        ptr = np.zeros(self.args.nE + 1, dtype=np.int64)
        nbr = np.zeros(max(tot.value, 1), dtype=np.int32)
        cls = np.zeros(max(tot.value, 1), dtype=np.uint8)
        check(L.nlfem_plan_neighbours(self._h, ptr.ctypes.data_as(C.POINTER(C.c_longlong)),
                                      nbr.ctypes.data_as(C.POINTER(C.c_int)),
                                      cls.ctypes.data_as(C.POINTER(C.c_ubyte))))
        return ptr, nbr[:tot.value], cls[:tot.value]

    STAT_KEYS = ("pairs", "probes", "close", "interior", "cut", "bary", "nnz_nodes", "nnz", "cap", "max_list", "subtri",
                 "ms_dual", "ms_bfs", "ms_transpose", "ms_pattern", "ms_positions", "ms_close", "ms_interior",
                 "ms_cut", "ms_compact", "ms_gather", "sur_pairs", "sur_rejected", "sur_items")

    def stats(self):
        out = np.zeros(len(self.STAT_KEYS))
        check(lib().nlfem_plan_stats(self._h, dp(out), out.size))
        return dict(zip(self.STAT_KEYS, out.tolist()))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().nlfem_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def stiffnessMatrix(mesh, kernel, configuration):
    """Stiffness matrix of the nonlocal operator as scipy CSR (cython/nlfem.pyx:128-274)."""
    verbose = configuration.get("verbose", 0)
    if "savePath" not in configuration:
        start = time.time()
        asm = Assembler(mesh, kernel, configuration)
        try:
            A = asm.assemble().csr()
        finally:
            asm.close()
        if verbose:
            print("Assembly Time\t", "{:1.2e}".format(time.time() - start), " Sec")
        return A
    a = parse(mesh, kernel, configuration)
    path = configuration["savePath"]
    Pg = a.Pg if a.Pg.size else np.zeros(1)
    dg = a.dg if a.dg.size else np.zeros(1)
    zi, zp, nz = _zeta(mesh)
    start = time.time()
    check(lib().nlfem_par_assemble(
        *_strs("system", path, ""), a.K_Omega, a.K, lp(a.elements), lp(a.elementLabels), dp(a.vertices),
        lp(a.vertexLabels), a.nE, a.nE_Omega, a.nV, a.nV_Omega, dp(a.Px), a.Px.size // 2, dp(a.dx), dp(a.Py),
        a.Py.size // 2, dp(a.dy), a.sqdelta, a.is_DG, 0, *_strs(a.kernel, ""),
        *_strs(a.method_remote, a.method_close), a.is_PlacePointOnCap, a.dim, a.outdim, lp(zi), lp(zp), nz,
        dp(Pg), a.degree, dp(dg), a.maxDiameter, a.fractional_s, a.is_fullConnectedComponentSearch, verbose))
    if verbose:
        print("Assembly Time\t", "{:1.2e}".format(time.time() - start), " Sec")
    return read_arma_spMat(path)


def loadVector(mesh, load, configuration):
    """Load vector f (cython/nlfem.pyx:276-389), assembled on the GPU."""
    kern = {"function": "constant", "horizon": 1.0, "outputdim": mesh.get("outdim", 1)}
    a = parse(mesh, kern, configuration)
    fd = np.zeros(a.K)
    zi, zp, nz = _zeta(mesh)
    check(lib().nlfem_load_vector(a.K, lp(a.elements), lp(a.elementLabels), dp(a.vertices), lp(a.vertexLabels),
                                  a.nE, a.nV, dp(a.Px), a.Px.size // 2, dp(a.dx), a.is_DG,
                                  load["function"].encode("utf-8"), a.outdim, dp(fd), lp(zi), lp(zp), nz))
    return fd


def evaluateMass(mesh, ud, Px, dx, J=None):
    """v = M u (cython/nlfem.pyx:638-676).  The reference passes J = nE_Omega, so only the
    first nE_Omega elements count; J=None keeps that behaviour, J=nE covers every element."""
    get = mesh.get if isinstance(mesh, dict) else (lambda k, d=None: getattr(mesh, k, d))
    el = np.ascontiguousarray(np.asarray(get("elements")), dtype=np.int64)
    lab = np.ascontiguousarray(np.asarray(get("elementLabels")), dtype=np.int64)
    vt = np.ascontiguousarray(np.asarray(get("vertices")), dtype=np.float64)
    vlab = np.ascontiguousarray(np.asarray(get("vertexLabels")), dtype=np.int64)
    is_DG = int(get("is_DiscontinuousGalerkin", 0))
    od = int(get("outdim", 1))
    nE, nV = el.shape[0], vt.shape[0]
    K = 3 * nE * od if is_DG else nV * od
    Px = np.ascontiguousarray(Px, dtype=np.float64).ravel()
    dx = np.ascontiguousarray(dx, dtype=np.float64).ravel()
    ud = np.ascontiguousarray(ud, dtype=np.float64).ravel()
    vd = np.zeros(K)
    if J is None:
        J = int(np.sum(lab > 0))
    check(lib().nlfem_par_evaluateMass(dp(vd), dp(ud), lp(el), lp(lab), dp(vt), lp(vlab), 0, int(J), dx.size,
                                       dp(Px), dp(dx), 2, od, is_DG, nE, nV))
    return vd


def get_vertexLabel(elements, elementLabels, vertices):
    """cython/nlfem.pyx:392-402: a vertex takes the smallest nonzero label of its elements."""
    nV = vertices.shape[0]
    vertexLabels = np.zeros(nV, dtype=np.int64)
    for label in np.sort(np.unique(elementLabels))[::-1]:
        if label:
            idx = np.where(elementLabels == label)
            vertexLabels[np.unique(elements[idx].ravel())] = label
    return vertexLabels


def get_diam(elements, vertices):
    """cython/nlfem.pyx:404-415 (keeps the reference's first-coordinate distance)."""
    T = vertices[elements]
    d = np.abs(T - np.roll(T, -1, axis=1))[:, :, 0]
    return float(d.max()) if d.size else 0.0


def meshFromArrays(elements, elementLabels, vertices, outputdim=1):
    """cython/nlfem.pyx:417-433"""
    vertexLabels = get_vertexLabel(elements, elementLabels, vertices)
    m = mesh_dict(np.asarray(elements, dtype=np.int64), elementLabels, np.asarray(vertices), vertexLabels,
                  outdim=outputdim, diam=get_diam(elements, vertices))
    return m


def setSolutionLabels(mesh, ansatz="CG", outputdim=1):
    """cython/nlfem.pyx:435-446"""
    mesh["ansatz"] = ansatz
    mesh["outputdim"] = outputdim
    if ansatz == "CG":
        mesh["solutionLabels"] = np.repeat(mesh["vertexLabels"], outputdim)
        mesh["K"] = mesh["nV"] * outputdim
        mesh["K_Omega"] = mesh["nV_Omega"] * outputdim
    if ansatz == "DG":
        mesh["solutionLabels"] = np.repeat(mesh["elementLabels"], 3 * outputdim)
        mesh["K"] = 3 * mesh["nE"] * outputdim
        mesh["K_Omega"] = 3 * mesh["nE_Omega"] * outputdim


def stiffnessMatrix_fromArray(elements, elementLabels, vertices, kernel, configuration):
    """cython/nlfem.pyx:449-481"""
    mesh = meshFromArrays(elements, elementLabels, vertices, kernel["outputdim"])
    setSolutionLabels(mesh, configuration["ansatz"], kernel["outputdim"])
    A = stiffnessMatrix(mesh, kernel, configuration)
    return mesh, A


def py_retriangulate(x_center, TE, delta, is_placePointOnCaps, pp=None):
    """cython/nlfem.pyx:712-726: (number of triangles, 9*3*2 list)."""
    x = np.ascontiguousarray(x_center, dtype=np.float64)
    T = np.ascontiguousarray(TE, dtype=np.float64)
    out = np.zeros(9 * 3 * 2)
    r = lib().nlfem_method_retriangulate(dp(x), dp(T), float(delta) ** 2, dp(out), int(is_placePointOnCaps))
    if r < 0:
        check(-r)
    return r, out


def retriangulate_batch(x, TE, sqdelta, caps=1):
    """Batched method_retriangulate on the GPU: x[n,2], TE[n,6] -> (ntri[n], lists[n,54])."""
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 2)
    T = np.ascontiguousarray(TE, dtype=np.float64).reshape(-1, 6)
    n = x.shape[0]
    lists = np.zeros((n, 54))
    ntri = np.zeros(n, dtype=np.int32)
    check(lib().nlfem_retriangulate_batch(n, dp(x), dp(T), float(sqdelta), int(caps), dp(lists),
                                          ntri.ctypes.data_as(C.POINTER(C.c_int))))
    return ntri, lists


def py_toRef(TE, phys_x):
    out = np.zeros(2)
    check(lib().nlfem_toRef(dp(np.ascontiguousarray(TE, dtype=np.float64)),
                            dp(np.ascontiguousarray(phys_x, dtype=np.float64)), dp(out)))
    return out


def py_toPhys(TE, p):
    out = np.zeros(2)
    check(lib().nlfem_toPhys(dp(np.ascontiguousarray(TE, dtype=np.float64)),
                             dp(np.ascontiguousarray(p, dtype=np.float64)), 2, dp(out)))
    return out


def py_solve2x2(A, b):
    out = np.zeros(2)
    check(lib().nlfem_solve2x2(dp(np.ascontiguousarray(A, dtype=np.float64)),
                               dp(np.ascontiguousarray(b, dtype=np.float64)), dp(out)))
    return out


__all__ = ["stiffnessMatrix", "loadVector", "evaluateMass", "read_arma_spMat", "read_arma_mat", "tensorgauss",
           "meshFromArrays", "get_vertexLabel", "setSolutionLabels", "stiffnessMatrix_fromArray", "py_retriangulate",
           "py_toRef", "py_toPhys", "py_solve2x2", "Assembler", "retriangulate_batch", "NlfemError"]
