"""The boundary as compiled callers see it: the Cython binding that INTEGRATION.md gives to
maintainers (tests/native/Cassemble_amd.pxd, nlfem_amd_binding.pyx) and a C++ caller of the
std::string drop-in par_assemble of include/Cassemble.h (tests/native/dropin_main.cpp, built as
the program dropin_main and as libdropin.so), all built by tests/native/build.py (from
__graft_entry__.build()).

CPU: all build and link against libnlfem_amd.so; an invalid configuration (dim = 3) is
reported as the reference's behaviour dictates: a RuntimeError with the message through the
Cython binding (status code), abort() with the message through the C++ drop-in (the program,
in a child process).
GPU: the binding and the C++ caller (libdropin.so, called in this process: the test starts no
child process once the GPU is in use) write the reference's Armadillo file for the C1 mesh, and
the file read back equals the oracle's matrix (pattern identical, entries within 1e-12 of the
row scale).
"""
import ctypes
import importlib
import os
import subprocess
import sys

import numpy as np
import pytest

from helpers import case, compare

HERE = os.path.dirname(os.path.abspath(__file__))
NATIVE = os.path.join(HERE, "native", "_build")
DROPIN = os.path.join(NATIVE, "dropin_main")
DROPIN_LIB = os.path.join(NATIVE, "libdropin.so")


def _built(gpu):
    have = os.path.exists(DROPIN) and os.path.exists(DROPIN_LIB) and any(f.startswith("nlfem_amd_binding") and f.endswith(".so")
                                          for f in os.listdir(NATIVE)) if os.path.isdir(NATIVE) else False
    if not have:
        if gpu:
            pytest.fail("tests/native is not built: run __graft_entry__.build() (or tests/native/build.py) first")
        sys.path.insert(0, os.path.join(HERE, "native"))
        import build as native_build
        native_build.build()
    if NATIVE not in sys.path:
        sys.path.insert(0, NATIVE)
    return importlib.import_module("nlfem_amd_binding")


def _args(mesh, conf, dim=2):
    from nlfem_amd.config import parse
    a = parse(mesh, {"function": "constant", "horizon": 0.1, "outputdim": 1}, conf)
    return a, dict(K_Omega=a.K_Omega, K=a.K, elements=a.elements, elementLabels=a.elementLabels,
                   vertices=a.vertices, vertexLabels=a.vertexLabels, nE=a.nE, nE_Omega=a.nE_Omega, nV=a.nV,
                   nV_Omega=a.nV_Omega, Px=a.Px, dx=a.dx, Py=a.Py, dy=a.dy, sqdelta=a.sqdelta, is_DG=0,
                   kernel="constant", method_remote="retriangulate", method_close="retriangulate", caps=1, dim=dim,
                   outdim=1)


def _write_input(path, a, dim=2):
    hdr = np.array([a.nE, a.nV, a.Px.size // 2, a.Py.size // 2, a.outdim, a.is_PlacePointOnCap, a.K, a.nE_Omega,
                    a.nV_Omega, dim], dtype=np.int64)
    with open(path, "wb") as f:
        for arr in (hdr, np.array([a.sqdelta]), a.elements, a.elementLabels, a.vertices, a.vertexLabels, a.Px, a.dx,
                    a.Py, a.dy):
            f.write(np.ascontiguousarray(arr).tobytes())


def test_binding_and_dropin_build_and_report_errors(tmp_path):
    mod = _built(gpu=False)
    mesh, kernel, conf = case("C1")
    a, kw = _args(mesh, conf, dim=3)
    kw["path"] = str(tmp_path / "A.arma")
    with pytest.raises(RuntimeError, match="dim"):
        mod.par_assemble_system(**kw)
    inp = tmp_path / "in.bin"
    _write_input(inp, a, dim=3)
    r = subprocess.run([DROPIN, str(inp), str(tmp_path / "B.arma"), "constant", "retriangulate", "retriangulate"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "dim" in r.stderr, (r.returncode, r.stderr)


@pytest.mark.gpu
def test_binding_and_dropin_match_oracle(tmp_path):
    import oracle
    from nlfem_amd import nlfem
    mod = _built(gpu=True)
    mesh, kernel, conf = case("C1")
    ref = oracle.par_system(mesh, kernel, conf)["A"]
    a, kw = _args(mesh, conf)
    kw["path"] = str(tmp_path / "A.arma")
    print("[native] Cython binding, C1", flush=True)
    assert mod.par_assemble_system(**kw) == 0
    A = nlfem.read_arma_spMat(kw["path"])
    same, rel, rowrel = compare(A, ref)
    assert same and rowrel <= 1e-12, (same, rowrel)
    print("[native] C++ std::string drop-in, C1", flush=True)
    inp = tmp_path / "in.bin"
    _write_input(inp, a)
    out = tmp_path / "B.arma"
    lib = ctypes.CDLL(DROPIN_LIB)
    lib.dropin_run.restype = ctypes.c_int
    lib.dropin_run.argtypes = [ctypes.c_char_p] * 5
    rc = lib.dropin_run(str(inp).encode(), str(out).encode(), b"constant", b"retriangulate", b"retriangulate")
    assert rc == 0, rc
    B = nlfem.read_arma_spMat(str(out))
    same, rel, rowrel = compare(B, ref)
    assert same and rowrel <= 1e-12, (same, rowrel)
