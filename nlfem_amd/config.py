"""Reading the (mesh, kernel, configuration) dictionaries of the reference Python API.

Mirrors the argument preparation of ``stiffnessMatrix`` (cython/nlfem.pyx:128-240):
dictionary keys, defaults and the K formula.  Kept quirk: the wrapper reads the cap switch
from ``approxBalls["is_PlacePointOnCap"]`` (default 1) while the shipped configurations spell
it ``isPlacePointOnCap``, so caps are on unless the former key says otherwise
(cython/nlfem.pyx:172).
"""
from dataclasses import dataclass
import numpy as np

from .quadrature import tensorgauss

SINGULAR_KERNELS = ("fractional", "linearPrototypeMicroelastic", "linearPrototypeMicroelasticField")
KERNELS = ("constant", "parabola", "fractional", "linearPrototypeMicroelastic",
           "linearPrototypeMicroelasticField")
METHODS = ("retriangulate", "baryCenter", "fractional", "weakSingular")


@dataclass
class AssemblyArgs:
    K_Omega: int
    K: int
    elements: np.ndarray
    elementLabels: np.ndarray
    vertices: np.ndarray
    vertexLabels: np.ndarray
    nE: int
    nE_Omega: int
    nV: int
    nV_Omega: int
    Px: np.ndarray
    dx: np.ndarray
    Py: np.ndarray
    dy: np.ndarray
    sqdelta: float
    is_DG: int
    kernel: str
    method_remote: str
    method_close: str
    is_PlacePointOnCap: int
    dim: int
    outdim: int
    Pg: np.ndarray
    degree: int
    dg: np.ndarray
    maxDiameter: float
    fractional_s: float
    is_fullConnectedComponentSearch: int
    verbose: int


def _c64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel())


def _ci64(a):
    return np.ascontiguousarray(np.asarray(a).astype(np.int64).ravel())


def parse(mesh, kernel, configuration):
    """Translate the three dictionaries into the flat argument list of ``par_assemble``."""
    dim = mesh["elements"].shape[1] - 1
    nE = mesh["elements"].shape[0]
    nV = mesh["vertices"].shape[0]
    nE_Omega = int(np.sum(mesh["elementLabels"] > 0))
    nV_Omega = int(np.sum(mesh["vertexLabels"] > 0))
    approx = configuration["approxBalls"]
    method_remote = approx["method"]
    method_close = configuration.get("closeElements", approx["method"])
    caps = int(approx.get("is_PlacePointOnCap", 1))
    quad = configuration["quadrature"]
    Px = _c64(quad["outer"]["points"])
    Py = _c64(quad["inner"]["points"])
    dx = _c64(quad["outer"]["weights"])
    dy = _c64(quad["inner"]["weights"])
    degree = int(quad.get("tensorGaussDegree", 0))
    if degree != 0:
        tg = tensorgauss(degree)
        Pg = _c64(tg.points)
        dg = _c64(tg.weights)
    else:
        Pg = np.zeros(0)
        dg = np.zeros(0)
    outdim = int(kernel["outputdim"])
    if mesh.get("outdim", outdim) != kernel["outputdim"]:
        raise ValueError("The output dimension of the mesh has to be equal to the output dimension of the kernel.")
    fractional_s = float(kernel.get("fractional_s", -1.0))
    is_DG = configuration.get("ansatz", "CG") == "DG"
    if is_DG:
        K = nE * (dim + 1) * outdim
        K_Omega = nE_Omega * (dim + 1) * outdim
    else:
        K = nV * outdim
        K_Omega = nV_Omega * outdim
    return AssemblyArgs(
        K_Omega=K_Omega, K=K,
        elements=_ci64(mesh["elements"]), elementLabels=_ci64(mesh["elementLabels"]),
        vertices=_c64(mesh["vertices"]), vertexLabels=_ci64(mesh["vertexLabels"]),
        nE=nE, nE_Omega=nE_Omega, nV=nV, nV_Omega=nV_Omega,
        Px=Px, dx=dx, Py=Py, dy=dy, sqdelta=float(kernel["horizon"]) ** 2,
        is_DG=int(is_DG), kernel=kernel["function"], method_remote=method_remote,
        method_close=method_close, is_PlacePointOnCap=caps, dim=dim, outdim=outdim,
        Pg=Pg, degree=degree, dg=dg, maxDiameter=float(mesh.get("diam", 0.0)),
        fractional_s=fractional_s,
        is_fullConnectedComponentSearch=int(configuration.get("is_fullConnectedComponentSearch", 0)),
        verbose=int(configuration.get("verbose", 0)),
    )


def make_configuration(method="retriangulate", close=None, outer="7point", inner="7point",
                       tensorGaussDegree=0, ansatz="CG", caps=1, precision="fp64", surrogate=None):
    """Build a configuration dictionary in the reference's layout (examples/test2D/conf/*.py)."""
    from .quadrature import rule
    Px, dx = rule(outer)
    Py, dy = rule(inner)
    conf = {
        "ansatz": ansatz,
        "approxBalls": {"method": method, "isPlacePointOnCap": bool(caps), "is_PlacePointOnCap": int(caps)},
        "quadrature": {
            "outer": {"points": Px, "weights": dx},
            "inner": {"points": Py, "weights": dy},
            "tensorGaussDegree": tensorGaussDegree,
        },
        "precision": precision,
    }
    if close is not None:
        conf["closeElements"] = close
    if surrogate is not None:
        conf["surrogate"] = surrogate
    return conf
