"""Manufactured-solution convergence harness (nlfem_amd/rates.py).

CPU: the harness itself, driven by the FP64 oracle (assembly, load, mass), reproduces the
P1 rate 2 of the reference's rates study; interpolation matches a brute-force barycentric
evaluation.  GPU: the GPU FP64 path gives the oracle's L2 errors to 1e-7; mixed precision
leaves errors within 0.1 % and rates within 0.005; the surrogate (active on the finer levels)
within 5 % and 0.05.
"""
import numpy as np
import pytest

import oracle
from nlfem_amd import rates


def _oracle_fns():
    def assemble(mesh, kernel, conf):
        return oracle.par_system(mesh, kernel, conf)["A"]

    def load(mesh, conf):
        return oracle.load_vector(mesh, {"function": "linear"}, conf)

    def mass(mesh, e, conf, J):
        return oracle.evaluate_mass(mesh, e, conf["quadrature"]["outer"]["points"],
                                    conf["quadrature"]["outer"]["weights"], J=J)
    return assemble, load, mass


def test_interpolation_matches_brute_force():
    n, delta = 6, 0.1
    m = rates.rates_mesh(n, delta)
    u = np.random.default_rng(0).normal(size=m["nV"])
    pts = np.random.default_rng(1).uniform(-delta, 0.5 + delta, size=(300, 2))
    got = rates.interpolate_p1(n, delta, u, pts)
    E = m["vertices"][m["elements"]]
    for p, g in zip(pts, got):
        for el, T in zip(m["elements"], E):
            M = np.array([T[1] - T[0], T[2] - T[0]]).T
            lam = np.linalg.solve(M, p - T[0])
            if lam.min() >= -1e-12 and lam.sum() <= 1 + 1e-12:
                ref = (1 - lam.sum()) * u[el[0]] + lam[0] * u[el[1]] + lam[1] * u[el[2]]
                assert abs(ref - g) < 1e-12
                break


def test_oracle_rates_are_two():
    a, l, m = _oracle_fns()
    r = rates.run(levels=(14, 28), delta=0.1, assemble=a, load=l, mass=m)
    print(f"[rates] oracle {r}", flush=True)
    assert 1.7 < r["rates"][1] < 2.4


@pytest.mark.gpu
def test_gpu_rates_match_oracle_and_are_unchanged():
    from nlfem_amd import nlfem
    levels = (28, 56, 112)  # the surrogate's domain covers the two finer levels
    a, l, m = _oracle_fns()
    ref = rates.run(levels=levels, delta=0.1, assemble=a, load=l, mass=m)
    gpu = rates.run(levels=levels, delta=0.1)
    print(f"[rates] oracle L2 {ref['L2']} rates {ref['rates']}", flush=True)
    print(f"[rates] gpu    L2 {gpu['L2']} rates {gpu['rates']}", flush=True)
    assert np.allclose(gpu["L2"], ref["L2"], rtol=1e-7, atol=0)

    def mixed(mesh, kernel, conf):
        c = dict(conf)
        c["precision"] = "mixed"
        return nlfem.stiffnessMatrix(mesh, kernel, c)

    def sur(mesh, kernel, conf):
        c = dict(conf)
        c["precision"] = "mixed"
        c["surrogate"] = {"model": "default"}
        asm = nlfem.Assembler(mesh, kernel, c)
        try:
            return asm.assemble().csr()
        finally:
            asm.close()

    # mixed precision: L2 errors within 0.1 %, rates within 0.005; surrogate (matrix error about
    # 1e-5 relative, on the two finer levels): L2 within 5 %, rates within 0.05
    for name, fn, l2tol, rtol in (("mixed", mixed, 1e-3, 0.005), ("surrogate+mixed", sur, 0.05, 0.05)):
        r = rates.run(levels=levels, delta=0.1, assemble=fn)
        print(f"[rates] {name} L2 {r['L2']} rates {r['rates']}", flush=True)
        assert np.allclose(r["L2"], ref["L2"], rtol=l2tol, atol=0)
        assert np.all(np.abs(np.array(r["rates"][1:]) - np.array(ref["rates"][1:])) < rtol)
