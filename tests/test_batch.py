"""Batched assembly of several meshes with different horizons in one plan (nlfem_amd/batch.py).

CPU: scale invariance of the kernels (the oracle's matrix of a mesh scaled by 1/delta with
horizon 1 equals the matrix of the mesh with horizon delta) and the stacking bookkeeping.
GPU: every diagonal block of the batched matrix against the oracle on its own mesh: pattern
identical, entries within 1e-12 of the sum of |addends| plus 1e-15 of the largest entry (the
rescaling perturbs coordinates by rounding, which tiny slivers of nearly tangent cuts amplify).
"""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.linalg import norm as spnorm

import oracle
from helpers import check_floor_usage, floor_usage
from nlfem_amd import batch
from nlfem_amd.config import make_configuration
from nlfem_amd.mesh import unit_square

FIELD = {"function": "linearPrototypeMicroelasticField", "outputdim": 2, "fractional_s": -0.5}


def _conf():
    return make_configuration(close="retriangulate", outer="16point", inner="7point")


def _meshes():
    return [(unit_square(10, 0.15, jitter=0.2, seed=1, outdim=2), 0.15),
            (unit_square(12, 0.1, jitter=0.2, seed=2, outdim=2), 0.1),
            (unit_square(8, 0.2, jitter=0.2, seed=3, outdim=2), 0.2)]


@pytest.mark.parametrize("kern", ["constant", "linearPrototypeMicroelasticField"])
def test_scale_invariance_on_oracle(kern):
    d = 0.15
    od = 2 if kern.endswith("Field") else 1
    m = unit_square(8, d, jitter=0.2, seed=5, outdim=od)
    k = {"function": kern, "horizon": d, "outputdim": od, "fractional_s": -0.5}
    conf = _conf() if od == 2 else make_configuration()
    A = oracle.par_system(m, k, conf)["A"]
    ms = dict(m)
    ms["vertices"] = m["vertices"] / d
    k1 = dict(k)
    k1["horizon"] = 1.0
    B = oracle.par_system(ms, k1, conf)["A"]
    # equal to rounding: 1e-12 of the sum of |addends| (the field kernel's identical pairs sum
    # near-singular addends ~1e6 times the entries) plus 1e-15 of the largest entry (slivers
    # of nearly tangent cuts amplify the rounding of the rescaled coordinates)
    Abs = oracle.abs_matrix(m, k, conf).tocsr()
    D = (A - B).tocoo()
    assert np.all(np.abs(D.data) <= 1e-12 * np.asarray(Abs[D.row, D.col]).ravel() + 1e-15 * abs(A).max())


def test_stack_offsets():
    ms = _meshes()
    u, voff, eoff = batch.stack([m for m, _ in ms], [d for _, d in ms])
    assert u["nE"] == sum(m["nE"] for m, _ in ms) and u["nV"] == sum(m["nV"] for m, _ in ms)
    assert np.array_equal(voff, np.cumsum([0] + [m["nV"] for m, _ in ms]))
    m1, d1 = ms[1]
    assert np.array_equal(u["elements"][eoff[1]:eoff[2]] - voff[1], m1["elements"])
    assert np.allclose(u["vertices"][voff[1]:voff[2]] * d1, m1["vertices"], rtol=1e-15, atol=0)


@pytest.mark.gpu
def test_batch_blocks_match_oracle():
    ms = _meshes()
    conf = _conf()
    b = batch.BatchAssembler([m for m, _ in ms], [d for _, d in ms], FIELD, conf)
    try:
        b.assemble()
        blocks = b.matrices()
        st = b.stats()
    finally:
        b.close()
    total = 0
    for (m, d), A in zip(ms, blocks):
        k = dict(FIELD)
        k["horizon"] = d
        ref = oracle.par_system(m, k, conf)
        total += ref["n_pairs"]
        R = ref["A"].tocsr()
        R.sort_indices()
        assert np.array_equal(A.indptr, R.indptr) and np.array_equal(A.indices, R.indices)
        Abs = oracle.abs_matrix(m, k, conf).tocsr()
        D = (A - R).tocoo()
        addends = np.asarray(Abs[D.row, D.col]).ravel()
        floor = 1e-15 * abs(R).max()
        bound = 1e-12 * addends + floor
        worst = float(np.max(np.abs(D.data) / np.maximum(bound, 1e-300))) if D.nnz else 0.0
        n_floor, worst_floorless = floor_usage(D.data, addends, 1e-12, floor)
        print(f"[batch] nE={m['nE']} delta={d} worst |diff|/bound = {worst:.3e}; entries needing the floor: "
              f"{n_floor} of {A.nnz} (worst floorless {worst_floorless:.3e})", flush=True)
        assert np.all(np.abs(D.data) <= bound)
        check_floor_usage(f"batch nE={m['nE']}", n_floor, worst_floorless, A.nnz)
    assert st["pairs"] == total


def _uneven():
    """Eight meshes of uneven size and horizon, as in the batched workload: h from 1/8 to 1/32,
    delta from below h (cut-dominated) to 6.4 h, vertex jitter 0.2 h."""
    spec = [(8, 0.05), (8, 0.2), (16, 0.1), (16, 0.05), (24, 0.15), (32, 0.2), (32, 0.06), (12, 0.12)]
    return [(unit_square(n, d, jitter=0.2, seed=20 + i, outdim=2), d) for i, (n, d) in enumerate(spec)]


@pytest.mark.gpu
def test_uneven_batch_blocks_match_oracle():
    ms = _uneven()
    conf = _conf()
    b = batch.BatchAssembler([m for m, _ in ms], [d for _, d in ms], FIELD, conf)
    try:
        b.assemble()
        blocks = b.matrices()
        st = b.stats()
    finally:
        b.close()
    total = 0
    for (m, d), A in zip(ms, blocks):
        k = dict(FIELD)
        k["horizon"] = d
        ref = oracle.par_system(m, k, conf)
        total += ref["n_pairs"]
        R = ref["A"].tocsr()
        R.sort_indices()
        assert np.array_equal(A.indptr, R.indptr) and np.array_equal(A.indices, R.indices)
        Abs = oracle.abs_matrix(m, k, conf).tocsr()
        D = (A - R).tocoo()
        addends = np.asarray(Abs[D.row, D.col]).ravel()
        floor = 1e-15 * abs(R).max()
        n_floor, worst_floorless = floor_usage(D.data, addends, 1e-12, floor)
        rel = float(spnorm(A - R) / spnorm(R))
        print(f"[uneven batch] nE={m['nE']} delta={d} rel fro {rel:.2e}; entries needing the floor: {n_floor} "
              f"of {A.nnz} (worst floorless {worst_floorless:.3e})", flush=True)
        assert np.all(np.abs(D.data) <= 1e-12 * addends + floor)
        check_floor_usage(f"uneven batch nE={m['nE']}", n_floor, worst_floorless, A.nnz)
    assert st["pairs"] == total
