"""Quadrature rules used by the assembly.

* ``P7``/``W7``: the 7-point degree-5 rule on the reference triangle with the 14-digit
  literals of the reference configurations (examples/test2D/conf/testConfConstant.py:20-33).
  The weights sum to 0.50000000000001 (the literal rounding is kept on purpose).
* ``P16``/``W16``: the 16-point rule with 8-digit points
  (examples/test2D/conf/testConfPeridynamics.py:38-70).
* :class:`tensorgauss`: deg**4-point tensor Gauss-Legendre rule on [-1,1]^4 with numpy's
  ``leggauss`` nodes, 'ij' meshgrid order, last index fastest (cython/nlfem.pyx:108-126).
"""
import numpy as np

P7 = np.array([[0.33333333333333, 0.33333333333333],
               [0.47014206410511, 0.47014206410511],
               [0.47014206410511, 0.05971587178977],
               [0.05971587178977, 0.47014206410511],
               [0.10128650732346, 0.10128650732346],
               [0.10128650732346, 0.79742698535309],
               [0.79742698535309, 0.10128650732346]])
W7 = 0.5 * np.array([0.22500000000000,
                     0.13239415278851,
                     0.13239415278851,
                     0.13239415278851,
                     0.12593918054483,
                     0.12593918054483,
                     0.12593918054483])

P16 = np.array([[0.33333333, 0.33333333],
                [0.45929259, 0.45929259],
                [0.45929259, 0.08141482],
                [0.08141482, 0.45929259],
                [0.17056931, 0.17056931],
                [0.17056931, 0.65886138],
                [0.65886138, 0.17056931],
                [0.05054723, 0.05054723],
                [0.05054723, 0.89890554],
                [0.89890554, 0.05054723],
                [0.26311283, 0.72849239],
                [0.72849239, 0.00839478],
                [0.00839478, 0.26311283],
                [0.72849239, 0.26311283],
                [0.26311283, 0.00839478],
                [0.00839478, 0.72849239]])
W16 = 0.5 * np.array([0.14431560767779, 0.09509163426728, 0.09509163426728, 0.09509163426728,
                      0.10321737053472, 0.10321737053472, 0.10321737053472, 0.03245849762320,
                      0.03245849762320, 0.03245849762320, 0.02723031417443, 0.02723031417443,
                      0.02723031417443, 0.02723031417443, 0.02723031417443, 0.02723031417443])

P1 = np.array([[1.0 / 3.0, 1.0 / 3.0]])
W1 = np.array([0.5])


class tensorgauss:
    """Tensor Gauss-Legendre rule with ``deg**dim`` points (cython/nlfem.pyx:108-126)."""

    def __init__(self, deg, dim=4):
        self.deg = deg
        self.dim = dim
        p, w = np.polynomial.legendre.leggauss(deg)
        self.N = deg ** dim
        idx = np.array([row.flatten() for row in
                        np.meshgrid(*([np.arange(deg)] * dim), indexing="ij")]).T
        self.points = p[idx].astype(np.float64)
        self.weights = np.ones(self.N)
        for j in range(dim):
            self.weights *= w[idx[:, j]]


def rule(name):
    """Return (points, weights) of a named triangle rule: '7point', '16point', '1point'."""
    if name in ("7point", "2d7p"):
        return P7.copy(), W7.copy()
    if name in ("16point", "2d16p"):
        return P16.copy(), W16.copy()
    if name == "1point":
        return P1.copy(), W1.copy()
    raise ValueError(f"unknown quadrature rule {name!r}")
