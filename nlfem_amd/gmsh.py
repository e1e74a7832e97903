"""gmsh 2.2 ASCII meshes with the reference tutorial's label conventions.

The reference's tutorial reads its meshes with meshio and joins every boundary line to one
"artificial" vertex (examples/tutorial/mesh.py:23-62):

- vertices are the nodes' first two coordinates, in the file's node order;
- elements are the triangles (element type 2), labelled with their physical tag; the largest
  tag (the nonlocal Dirichlet layer) is negated;
- with ``set_art_vertex`` a vertex at (0, 0) is appended and every line element (type 1)
  [a, b] becomes a triangle [a, b, artificial] with label 0.  Label-0 elements are never
  integrated, only traversed by the neighbour search, so the artificial vertex connects every
  boundary to every other one (the disjoint domains of ``disjoint.msh`` become one connected
  dual graph).

Only the ASCII variant of format 2.2 is read ($MeshFormat 2.2 0 8).
"""
import numpy as np


def _sections(path):
    with open(path, "r") as f:
        lines = [ln.strip() for ln in f]
    out, i = {}, 0
    while i < len(lines):
        ln = lines[i]
        if ln.startswith("$") and not ln.startswith("$End"):
            name = ln[1:]
            j = lines.index("$End" + name, i + 1)
            out[name] = lines[i + 1:j]
            i = j + 1
        else:
            i += 1
    return out


def read_msh22(path):
    """Nodes and elements of a gmsh 2.2 ASCII file.

    Returns (points [nN, 3] float64, {element type: (connectivity [n, k] int64 zero-based,
    physical tags [n] int64)}).  Node numbers are mapped to their position in the $Nodes
    section."""
    sec = _sections(path)
    fmt = sec.get("MeshFormat", [""])[0].split()
    if len(fmt) < 2 or not fmt[0].startswith("2.2") or fmt[1] != "0":
        raise ValueError(f"{path}: not a gmsh 2.2 ASCII mesh (format line {fmt})")
    nodes = sec["Nodes"]
    n = int(nodes[0])
    ids = np.empty(n, dtype=np.int64)
    pts = np.empty((n, 3), dtype=np.float64)
    for k in range(n):
        parts = nodes[1 + k].split()
        ids[k] = int(parts[0])
        pts[k] = [float(v) for v in parts[1:4]]
    index = {int(t): k for k, t in enumerate(ids)}
    elems = sec["Elements"]
    m = int(elems[0])
    nnode = {1: 2, 2: 3, 15: 1}
    conn, tags = {}, {}
    for k in range(m):
        parts = [int(v) for v in elems[1 + k].split()]
        etype, ntags = parts[1], parts[2]
        if etype not in nnode:
            raise ValueError(f"{path}: element type {etype} is not supported")
        phys = parts[3] if ntags > 0 else 0
        vs = parts[3 + ntags:3 + ntags + nnode[etype]]
        conn.setdefault(etype, []).append([index[v] for v in vs])
        tags.setdefault(etype, []).append(phys)
    cells = {t: (np.asarray(conn[t], dtype=np.int64), np.asarray(tags[t], dtype=np.int64)) for t in conn}
    return pts, cells


def read_gmsh(path, set_art_vertex=True):
    """(vertices [nV, 2], elements [nE, 3], elementLabels [nE]) with the tutorial's labels
    (examples/tutorial/mesh.py:51-62) and, by default, the artificial vertex
    (examples/tutorial/mesh.py:23-34)."""
    pts, cells = read_msh22(path)
    vertices = np.ascontiguousarray(pts[:, :2])
    elements, labels = cells[2]
    labels = labels.copy()
    labels[labels == labels.max()] *= -1
    if set_art_vertex:
        n_vertex = vertices.shape[0]
        vertices = np.concatenate([vertices, [[0.0, 0.0]]])
        lines = cells[1][0]
        new = np.concatenate([lines, np.full((lines.shape[0], 1), n_vertex, dtype=np.int64)], axis=1)
        elements = np.concatenate([elements, new])
        labels = np.concatenate([labels, np.zeros(new.shape[0], dtype=np.int64)])
    return vertices, np.ascontiguousarray(elements, dtype=np.int64), np.ascontiguousarray(labels, dtype=np.int64)
