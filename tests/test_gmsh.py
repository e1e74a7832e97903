"""The gmsh 2.2 reader and the tutorial's artificial-vertex construction (nlfem_amd/gmsh.py)
on the reference's own meshes (examples/tutorial/data/*.msh, fixtures under tests/golden)."""
import numpy as np
import pytest

from helpers import GOLDEN, gmsh_case
from nlfem_amd.gmsh import read_gmsh, read_msh22

# (nodes, triangles, lines) as counted in each file's $Nodes / $Elements sections
COUNTS = {"hole": (281, 466, 128), "disjoint": (524, 820, 224), "circle": (735, 1386, 146)}


@pytest.mark.parametrize("name", sorted(COUNTS))
def test_reader_counts_labels_and_artificial_vertex(name):
    import os
    path = os.path.join(GOLDEN, "meshes", f"{name}.msh")
    pts, cells = read_msh22(path)
    nN, nT, nL = COUNTS[name]
    assert pts.shape == (nN, 3) and cells[2][0].shape == (nT, 3) and cells[1][0].shape == (nL, 2)
    v0, e0, l0 = read_gmsh(path, set_art_vertex=False)
    assert v0.shape == (nN, 2) and e0.shape == (nT, 3) and not np.any(l0 == 0)
    # the largest physical tag (the nonlocal Dirichlet layer) is negated, the others stay positive
    tags = cells[2][1]
    assert np.array_equal(l0 < 0, tags == tags.max())
    v, e, lab = read_gmsh(path)
    assert v.shape == (nN + 1, 2) and np.array_equal(v[-1], [0.0, 0.0])
    assert e.shape == (nT + nL, 3) and np.all(lab[nT:] == 0) and np.all(e[nT:, 2] == nN)
    assert np.array_equal(e[nT:, :2], cells[1][0])
    # every triangle references existing nodes; the artificial vertex is shared by all lines
    assert e.min() >= 0 and e.max() == nN and np.bincount(e.ravel())[nN] == nL


@pytest.mark.parametrize("name", sorted(COUNTS))
def test_vertex_labels_follow_the_reference_rule(name):
    """get_vertexLabel (cython/nlfem.pyx:392-402): the artificial vertex (only label-0
    elements) keeps label 0; every other vertex takes the smallest nonzero label around it."""
    mesh, _, _ = gmsh_case(name)
    el, lab, vl = mesh["elements"], mesh["elementLabels"], mesh["vertexLabels"]
    assert vl[-1] == 0
    for v in range(0, mesh["nV"] - 1, 7):
        around = lab[np.any(el == v, axis=1)]
        nz = around[around != 0]
        assert vl[v] == (nz.min() if nz.size else 0)
