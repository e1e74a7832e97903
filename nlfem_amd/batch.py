"""Many meshes in one assembly: every phase of the pipeline runs as one launch over all of them.

Each mesh m with horizon delta_m is scaled by 1/delta_m.  Every kernel of the path is
normalised so that its second moment is 2 (reference src/model.cpp:25-60,116-153): constant
4/(pi delta^4), parabola 12/(pi delta^6) (delta^2 - |z|^2), fractional
(2-2s)/(pi delta^(2-2s)) |z|^-(2+2s), microelastic 3/(pi delta^3) / |z| and its field form
3/delta^3 z z^T / |z|^3.  Each of them satisfies k_delta(z) = delta^-4 k_1(z / delta), and the
element areas scale by delta^2 on each side of a pair, so the stiffness matrix of the scaled
mesh with horizon 1 equals the matrix of the original mesh with horizon delta_m.  The scaled
meshes are stacked into one disconnected mesh (vertex and element indices offset, no
translation: the BFS walks the dual graph, which has no edges between meshes, so meshes may
overlap in space).  One plan with horizon 1 then assembles all of them: the BFS, pattern and
quadrature kernels each run once over all outer elements.  The union matrix is block
diagonal; block m is the matrix of mesh m.

The values match the per-mesh assembly to rounding (the coordinates are rescaled), not bit
for bit.
"""
import numpy as np

from .mesh import mesh_dict
from ._lib import check, lib
from .nlfem import Assembler

SCALE_INVARIANT = ("constant", "parabola", "fractional", "linearPrototypeMicroelastic",
                   "linearPrototypeMicroelasticField")


def stack(meshes, deltas):
    """Union mesh of the meshes scaled by 1/delta; returns (mesh dict, vertex offsets, element offsets)."""
    if len(meshes) != len(deltas) or not meshes:
        raise ValueError("one delta per mesh")
    od = {m.get("outdim", 1) for m in meshes}
    an = {m.get("ansatz", "CG") for m in meshes}
    if len(od) != 1 or len(an) != 1:
        raise ValueError("meshes of one batch share outdim and ansatz")
    el, elab, vt, vlab = [], [], [], []
    voff, eoff = [0], [0]
    for m, d in zip(meshes, deltas):
        if not d > 0:
            raise ValueError("delta must be positive")
        el.append(np.asarray(m["elements"], dtype=np.int64) + voff[-1])
        elab.append(np.asarray(m["elementLabels"], dtype=np.int64))
        vt.append(np.asarray(m["vertices"], dtype=np.float64) / d)
        vlab.append(np.asarray(m["vertexLabels"], dtype=np.int64))
        voff.append(voff[-1] + m["nV"])
        eoff.append(eoff[-1] + m["nE"])
    u = mesh_dict(np.concatenate(el), np.concatenate(elab), np.concatenate(vt), np.concatenate(vlab),
                  outdim=od.pop(), ansatz=an.pop())
    return u, np.array(voff), np.array(eoff)


class BatchAssembler:
    """Assemble the stiffness matrices of several meshes (each with its own horizon) together."""

    def __init__(self, meshes, deltas, kernel, configuration):
        if kernel["function"] not in SCALE_INVARIANT:
            raise ValueError(f"kernel {kernel['function']} is not scale invariant")
        self.meshes = list(meshes)
        self.deltas = [float(d) for d in deltas]
        self.union, self.voff, self.eoff = stack(self.meshes, self.deltas)
        k = dict(kernel)
        k["horizon"] = 1.0
        self.asm = Assembler(self.union, k, configuration)
        # capacities from the mesh with the most elements per (unit) horizon
        best = None
        for m, d in zip(self.meshes, self.deltas):
            E = np.asarray(m["vertices"])[np.asarray(m["elements"])] / d
            edges = np.linalg.norm(E - np.roll(E, -1, axis=1), axis=2)
            area = 0.5 * np.abs((E[:, 1, 0] - E[:, 0, 0]) * (E[:, 2, 1] - E[:, 0, 1])
                                - (E[:, 2, 0] - E[:, 0, 0]) * (E[:, 1, 1] - E[:, 0, 1]))
            hmax, amin, amean = float(edges.max()), float(area.min()), float(area.mean())
            aref = amin if amin > 0.6 * amean else 0.6 * amean
            load = np.pi * (1.0 + 3.0 * hmax) ** 2 / aref
            if best is None or load > best[0]:
                best = (load, hmax, amin, amean)
        check(lib().nlfem_plan_set_estimates(self.asm._h, best[1], best[2], best[3]))

    def assemble(self):
        self.asm.assemble()
        return self

    def autotune(self):
        """Pairs per chunk of the cut kernel tuned on this batch (see Assembler.autotune)."""
        return self.asm.autotune()

    def set_surrogate(self, model="default", tol=None, smin=None, smax=None, enabled=True):
        """Cut pairs through the surrogate (see Assembler.set_surrogate; the batch's horizon is 1
        after scaling, which is what the networks' delta-normalised inputs expect)."""
        self.asm.set_surrogate(model, tol, smin, smax, enabled)
        return self

    def stats(self):
        return self.asm.stats()

    def matrices(self):
        """Per-mesh CSR matrices (diagonal blocks of the union matrix)."""
        A = self.asm.csr()
        out = []
        is_dg = self.union["is_DiscontinuousGalerkin"]
        od = self.union["outdim"]
        for m in range(len(self.meshes)):
            if is_dg:
                lo, hi = 3 * od * self.eoff[m], 3 * od * self.eoff[m + 1]
            else:
                lo, hi = od * self.voff[m], od * self.voff[m + 1]
            B = A[lo:hi, lo:hi].tocsr()
            B.sort_indices()
            out.append(B)
        return out

    def close(self):
        self.asm.close()


def c5_meshes(count=48, seed=0, max_ratio=12.0, outdim=2):
    """Batched uneven workload: h from {1/32, 1/64, 1/128, 1/256}, jitter +-0.2h, delta uniform in
    [0.05, 0.2] limited to delta <= max_ratio * h.  One plan gives every outer element the
    neighbour-list capacity of the mesh with the most elements per horizon; at delta = 0.2 on
    h = 1/256 (51 h) that is ~16,000 slots for each of the ~1.9 M elements (~120 GB for the
    lists, as much for their transpose), so the workload keeps the horizons at 12 h."""
    from .mesh import unit_square
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.choice([32, 64, 128, 256]))
        hi = min(0.2, max_ratio / n)
        d = float(rng.uniform(0.05, hi)) if hi > 0.05 else hi
        out.append((unit_square(n, d, jitter=0.2, seed=int(rng.integers(1 << 30)), outdim=outdim), d, n))
    return out
