"""Manufactured-solution convergence study on the GPU assembly.

Follows the reference's rates driver (examples/test2D/computeRates2D.py:17-110 with the
constant-kernel configuration examples/test2D/conf/testConfConstant.py): Omega = (0, 0.5)^2
inside the mesh of [-delta, 0.5 + delta]^2 (examples/test2D/mesh.py:24-44), exact solution
u(x) = x0^2 x1 + x1^2 with load f = -2 (x1 + 1) ("linear"), nonlocal Dirichlet data
g = u on the layer, CG on A_OO u = f_O - A_OI g (tol 1e-9), the discrete solution
interpolated to a mesh four times finer than the finest level, and the L2 error
sqrt(e^T M e) over Omega with the mass-matrix action (par_evaluateMass).  Levels use
n = n0 * 2^l squares per side.
"""
import numpy as np
import scipy.sparse.linalg as spla

from . import nlfem
from .config import make_configuration
from .mesh import unit_square

LO, HI = 0.0, 0.5


def u_exact(x):
    x = np.asarray(x, dtype=np.float64)
    return x[:, 0] ** 2 * x[:, 1] + x[:, 1] ** 2


def rates_mesh(n, delta):
    return unit_square(n, delta, lo=LO - delta, hi=HI + delta)


def interpolate_p1(n, delta, u, points):
    """Value at `points` of the P1 function with nodal values u on rates_mesh(n, delta)."""
    lo = LO - delta
    h = (HI + delta - lo) / n
    t = (np.asarray(points) - lo) / h
    i = np.clip(np.floor(t[:, 0] + 1e-12).astype(np.int64), 0, n - 1)
    j = np.clip(np.floor(t[:, 1] + 1e-12).astype(np.int64), 0, n - 1)
    a = t[:, 0] - i
    b = t[:, 1] - j
    v00 = j * (n + 1) + i
    v10, v01 = v00 + 1, v00 + n + 1
    v11 = v01 + 1
    lower = a >= b  # element [v00, v10, v11] below the "up" diagonal, [v00, v11, v01] above
    out = np.where(lower, (1 - a) * u[v00] + (a - b) * u[v10] + b * u[v11],
                   (1 - b) * u[v00] + (b - a) * u[v01] + a * u[v11])
    return out


def solve_level(mesh, A, f):
    """Nodal solution on all vertices: u = g on the layer, CG on the Omega block."""
    lab = mesh["nodeLabels"]
    om, ly = lab > 0, lab < 0
    g = u_exact(mesh["vertices"][ly])
    A = A.tocsr()
    rhs = f[om] - A[om][:, ly] @ g
    x, info = spla.cg(A[om][:, om], rhs, x0=rhs, rtol=1e-9, atol=0.0, maxiter=20000)
    if info != 0:
        raise RuntimeError(f"CG did not converge (info={info})")
    u = np.empty(mesh["nV"])
    u[om] = x
    u[ly] = g
    return u


def l2_error(n, delta, u, n_fine, conf, mass=None):
    """L2(Omega) norm of u_exact - I_h u on the fine mesh, through the mass action (GPU by default)."""
    fine = rates_mesh(n_fine, delta)
    order = np.argsort(fine["elementLabels"] <= 0, kind="stable")  # Omega elements first
    fm = dict(fine)
    fm["elements"] = fine["elements"][order]
    fm["elementLabels"] = fine["elementLabels"][order]
    e = u_exact(fine["vertices"]) - interpolate_p1(n, delta, u, fine["vertices"])
    J = int(np.sum(fine["elementLabels"] > 0))
    if mass is None:
        Me = nlfem.evaluateMass(fm, e, conf["quadrature"]["outer"]["points"], conf["quadrature"]["outer"]["weights"],
                                J=J)
    else:
        Me = mass(fm, e, conf, J)
    return float(np.sqrt(max(e @ Me, 0.0)))


def run(levels=(14, 28, 56), delta=0.1, assemble=None, load=None, mass=None, log=None):
    """L2 errors and rates.  assemble(mesh, kernel, conf) -> scipy CSR, load(mesh, conf) -> f and
    mass(mesh, e, conf, J) -> M e default to the GPU paths of this package."""
    kernel = {"function": "constant", "horizon": delta, "outputdim": 1}
    conf = make_configuration()
    if assemble is None:
        def assemble(mesh, kernel, conf):
            return nlfem.stiffnessMatrix(mesh, kernel, conf)
    n_fine = 4 * levels[-1]
    hs, errs, rates = [], [], []
    for n in levels:
        mesh = rates_mesh(n, delta)
        A = assemble(mesh, kernel, conf)
        f = nlfem.loadVector(mesh, {"function": "linear"}, conf) if load is None else load(mesh, conf)
        u = solve_level(mesh, A, f)
        err = l2_error(n, delta, u, n_fine, conf, mass)
        rates.append(float(np.log(errs[-1] / err) / np.log(2.0)) if errs else 0.0)
        hs.append((HI - LO + 2 * delta) / n)
        errs.append(err)
        if log:
            log(f"n={n} h={hs[-1]:.4e} L2={err:.6e} rate={rates[-1]:.3f}")
    return {"h": hs, "L2": errs, "rates": rates}
