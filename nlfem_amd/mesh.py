"""Seeded unit-square triangle meshes for the assembly.

The domain is the unit square [0,1]^2 cut into ``n x n`` squares, each split along its
"up" diagonal (lower-left to upper-right) into two counter-clockwise triangles, so
``nE = 2 n^2`` and ``nV = (n+1)^2``.  Vertex ``(i, j)`` has index ``j*(n+1) + i``; square
``(i, j)`` holds elements ``2*(j*n+i)`` = [v00, v10, v11] and ``2*(j*n+i)+1`` = [v00, v11, v01].

Labels follow the reference rates mesh (examples/test2D/mesh.py:179-196) with
Omega = (delta, 1-delta)^2 and the nonlocal Dirichlet layer Gamma = square minus Omega:
a vertex is +1 if it lies strictly inside Omega (tolerance 1e-9) on the unperturbed
lattice and -1 otherwise; an element is +1 if any of its vertices is +1, else -1.
No element has label 0.

Jitter displaces interior lattice vertices by U(-j*h, j*h) per coordinate from
``numpy.random.default_rng(seed)``; boundary vertices stay fixed.  Any vertex of a triangle
whose signed area drops below 10% of the unperturbed area is redrawn (same generator) until
every triangle is valid, so the mesh is a function of (n, jitter, seed) alone.
"""
import numpy as np


def unit_square(n, delta, jitter=0.0, seed=0, outdim=1, ansatz="CG", lo=0.0, hi=1.0):
    """Mesh of [lo, hi]^2 (default the unit square) with Omega = (lo+delta, hi-delta)^2."""
    h = (hi - lo) / n
    ii, jj = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="xy")
    lattice = lo + np.stack([ii.ravel() * h, jj.ravel() * h], axis=1).astype(np.float64)
    verts = lattice.copy()

    sq_i, sq_j = np.meshgrid(np.arange(n), np.arange(n), indexing="xy")
    sq_i = sq_i.ravel()
    sq_j = sq_j.ravel()
    v00 = sq_j * (n + 1) + sq_i
    v10 = v00 + 1
    v01 = v00 + (n + 1)
    v11 = v01 + 1
    elems = np.empty((2 * n * n, 3), dtype=np.int64)
    elems[0::2] = np.stack([v00, v10, v11], axis=1)
    elems[1::2] = np.stack([v00, v11, v01], axis=1)

    if jitter > 0.0:
        rng = np.random.default_rng(seed)
        interior = np.where((ii.ravel() > 0) & (ii.ravel() < n) & (jj.ravel() > 0) & (jj.ravel() < n))[0]
        amp = jitter * h
        verts[interior] = lattice[interior] + rng.uniform(-amp, amp, size=(interior.size, 2))
        min_area = 0.1 * 0.5 * h * h
        for _ in range(1000):
            E = verts[elems]
            area = 0.5 * ((E[:, 1, 0] - E[:, 0, 0]) * (E[:, 2, 1] - E[:, 0, 1])
                          - (E[:, 2, 0] - E[:, 0, 0]) * (E[:, 1, 1] - E[:, 0, 1]))
            bad = area < min_area
            if not bad.any():
                break
            redo = np.unique(elems[bad].ravel())
            redo = redo[np.isin(redo, interior)]
            verts[redo] = lattice[redo] + rng.uniform(-amp, amp, size=(redo.size, 2))
        else:
            raise RuntimeError("jitter resampling did not converge")

    inside = np.all((lattice > lo + delta + 1e-9) & (lattice < hi - delta - 1e-9), axis=1)
    vlabels = np.where(inside, 1, -1).astype(np.int64)
    elabels = np.where((vlabels[elems] > 0).any(axis=1), 1, -1).astype(np.int64)
    return mesh_dict(elems, elabels, verts, vlabels, outdim=outdim, ansatz=ansatz, diam=np.sqrt(2.0) * h * (1 + 2 * jitter))


def mesh_dict(elements, elementLabels, vertices, vertexLabels, outdim=1, ansatz="CG", diam=0.0):
    """Mesh dictionary with the keys the reference wrapper reads (cython/nlfem.pyx:157-167,417-433)."""
    elements = np.ascontiguousarray(elements, dtype=np.int64)
    elementLabels = np.ascontiguousarray(elementLabels, dtype=np.int64)
    vertices = np.ascontiguousarray(vertices, dtype=np.float64)
    vertexLabels = np.ascontiguousarray(vertexLabels, dtype=np.int64)
    nE, nV = elements.shape[0], vertices.shape[0]
    dim = elements.shape[1] - 1
    mesh = {
        "dim": dim,
        "nE": nE,
        "nV": nV,
        "nE_Omega": int(np.sum(elementLabels > 0)),
        "nV_Omega": int(np.sum(vertexLabels > 0)),
        "elements": elements,
        "elementLabels": elementLabels,
        "vertices": vertices,
        "vertexLabels": vertexLabels,
        "diam": diam,
        "outdim": outdim,
        "ansatz": ansatz,
    }
    if ansatz == "DG":
        mesh["K"] = 3 * nE * outdim
        mesh["K_Omega"] = 3 * mesh["nE_Omega"] * outdim
        mesh["is_DiscontinuousGalerkin"] = 1
        mesh["nodeLabels"] = np.repeat(elementLabels, 3 * outdim)
    else:
        mesh["K"] = nV * outdim
        mesh["K_Omega"] = mesh["nV_Omega"] * outdim
        mesh["is_DiscontinuousGalerkin"] = 0
        mesh["nodeLabels"] = np.repeat(vertexLabels, outdim)
    return mesh
