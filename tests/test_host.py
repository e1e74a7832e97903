"""Host-side checks without a GPU: C ABI exports, configuration parsing, mesh generator,
Armadillo file reader."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from nlfem_amd import _lib
from nlfem_amd.config import make_configuration, parse
from nlfem_amd.mesh import unit_square
from nlfem_amd.nlfem import read_arma_mat, read_arma_spMat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    header = open(os.path.join(ROOT, "include", "nlfem_amd.h")).read()
    names = set(re.findall(r"\b(nlfem_\w+)\s*\(", header))
    assert len(names) >= 18
    for n in names:
        assert hasattr(L, n), n
    assert set(_lib.SIGNATURES) <= names


def test_library_exports_cassemble_cpp_dropin():
    out = subprocess.run(["nm", "-D", "-C", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    for sym in ("par_assemble(std::", "par_evaluateMass(double*", "method_retriangulate(double const*",
                "method_retriangulateInfty(", "toRef(double const*", "toPhys(double const*, double const*, int",
                "solve2x2(double const*"):
        assert sym in out, sym


def test_error_reporting_without_device():
    """Invalid configurations are rejected on the host before any device call."""
    L = _lib.lib()
    h = C.c_void_p()
    el = np.array([0, 1, 2], dtype=np.int64)
    lab = np.array([1], dtype=np.int64)
    vt = np.array([0.0, 0.0, 1.0, 0.0, 0.0, 1.0])
    vl = np.ones(3, dtype=np.int64)
    P = np.array([1 / 3, 1 / 3])
    w = np.array([0.5])
    rc = L.nlfem_plan_create(C.byref(h), 3, _lib.lp(el), _lib.lp(lab), _lib.dp(vt), _lib.lp(vl), 1, 3, _lib.dp(P), 1,
                             _lib.dp(w), _lib.dp(P), 1, _lib.dp(w), 0.01, 0, b"nope", b"retriangulate",
                             b"retriangulate", 1, 1, None, 0, None, -1.0, 0, 0)
    assert rc == 1 and b"unknown kernel" in L.nlfem_last_error()
    rc = L.nlfem_plan_create(C.byref(h), 3, _lib.lp(el), _lib.lp(lab), _lib.dp(vt), _lib.lp(vl), 1, 3, _lib.dp(P), 1,
                             _lib.dp(w), _lib.dp(P), 1, _lib.dp(w), 0.01, 0, b"constant", b"fractional",
                             b"retriangulate", 1, 1, None, 0, None, -1.0, 0, 0)
    assert rc == 1 and b"close elements only" in L.nlfem_last_error()
    rc = L.nlfem_plan_create(C.byref(h), 3, _lib.lp(el), _lib.lp(lab), _lib.dp(vt), _lib.lp(vl), 1, 3, _lib.dp(P), 1,
                             _lib.dp(w), _lib.dp(P), 1, _lib.dp(w), 0.01, 0, b"linearPrototypeMicroelastic",
                             b"retriangulate", b"retriangulate", 1, 1, None, 0, None, -1.0, 0, 0)
    assert rc == 1 and b"s=-0.5" in L.nlfem_last_error()


def test_parse_defaults_and_quirks():
    mesh = unit_square(4, 0.2)
    kernel = {"function": "constant", "horizon": 0.2, "outputdim": 1}
    conf = make_configuration()
    conf["approxBalls"] = {"method": "retriangulate", "isPlacePointOnCap": False}
    a = parse(mesh, kernel, conf)
    assert a.is_PlacePointOnCap == 1  # the wrapper reads 'is_PlacePointOnCap' (default 1)
    assert a.method_close == "retriangulate" and a.degree == 0 and a.fractional_s == -1.0
    assert a.K == mesh["nV"] and a.sqdelta == pytest.approx(0.04)
    conf["ansatz"] = "DG"
    assert parse(mesh, kernel, conf).K == 3 * mesh["nE"]
    with pytest.raises(ValueError):
        parse(dict(mesh, outdim=2), kernel, conf)


def test_unit_square_mesh():
    m = unit_square(16, 0.1)
    assert m["nE"] == 512 and m["nV"] == 289
    E = m["vertices"][m["elements"]]
    area = 0.5 * ((E[:, 1, 0] - E[:, 0, 0]) * (E[:, 2, 1] - E[:, 0, 1]) - (E[:, 2, 0] - E[:, 0, 0]) * (E[:, 1, 1] - E[:, 0, 1]))
    assert np.allclose(area, 0.5 / 256)
    inside = (m["vertices"] > 0.1 + 1e-9) & (m["vertices"] < 0.9 - 1e-9)
    assert np.array_equal(m["vertexLabels"] > 0, inside.all(axis=1))
    j = unit_square(32, 0.1, jitter=0.3, seed=5)
    E = j["vertices"][j["elements"]]
    area = 0.5 * ((E[:, 1, 0] - E[:, 0, 0]) * (E[:, 2, 1] - E[:, 0, 1]) - (E[:, 2, 0] - E[:, 0, 0]) * (E[:, 1, 1] - E[:, 0, 1]))
    assert area.min() > 0.1 * 0.5 / 32 ** 2
    b = (j["vertices"] == 0) | (j["vertices"] == 1)
    assert np.array_equal(unit_square(32, 0.1, jitter=0.3, seed=5)["vertices"], j["vertices"])
    assert b.any(axis=1).sum() == 4 * 32


def test_read_arma_files(tmp_path):
    A = np.array([[1.0, 0, 2.0], [0, 3.0, 0], [4.0, 0, 5.0]])
    # file holds A^T in CSC = A in CSR: values, column indices, row pointers (uint64)
    vals = np.array([1.0, 2.0, 3.0, 4.0, 5.0])
    cols = np.array([0, 2, 1, 0, 2], dtype=np.uint64)
    rp = np.array([0, 2, 3, 5], dtype=np.uint64)
    p = tmp_path / "A.bin"
    with open(p, "wb") as f:
        f.write(b"ARMA_SPM_BIN_FN008\n3 3 5\n")
        f.write(vals.tobytes() + cols.tobytes() + rp.tobytes())
    assert np.array_equal(read_arma_spMat(str(p)).toarray(), A)
    q = tmp_path / "f.bin"
    with open(q, "wb") as f:
        f.write(b"ARMA_MAT_BIN_FN008\n3 1\n" + np.array([1.0, 2.0, 3.0]).tobytes())
    assert np.array_equal(read_arma_mat(str(q))[:, 0], [1.0, 2.0, 3.0])
    with open(q, "wb") as f:
        f.write(b"WRONG\n")
    with pytest.raises(ValueError):
        read_arma_mat(str(q))
