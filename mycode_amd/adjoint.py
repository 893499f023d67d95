"""Discrete-adjoint design gradients on the GPU (no reference counterpart).

For a converged state U*(alpha) with R(U*, X(alpha)) = 0 and a functional
J(U, X), the total derivative is

    dJ/dalpha_k = dJ/dX . dX/dalpha_k  -  psi^T (dR/dX . dX/dalpha_k),
    (dR/dU)^T psi = (dJ/dU)^T,

psi from GMRES on the transposed system (the library's reverse product,
``jv = 2``), and the two partials at fixed U by central differences of the
geometry (the residual and J are evaluated on the GPU with the moved mesh).

Configuration 4: J = sum over the bump faces (bottom wall, 1 <= x <= 2) of
p(U_owner) A_f; 64 design variables are equispaced hat functions on the bump
moving the wall nodes vertically, the displacement decaying linearly to zero
at the top wall.
"""
from __future__ import annotations

import numpy as np

from . import cfd


def bump_faces(faces) -> np.ndarray:
    return np.where((faces["neighbor"] < 0) & (faces["bc_tag"] == 3) & (faces["x_mid"] >= 1.0)
                    & (faces["x_mid"] <= 2.0) & (faces["y_mid"] < 0.5))[0]


def pressure(p: cfd.Params, U):
    return (p.gamma - 1.0) * (U[..., 3] - 0.5 * U[..., 0] * ((U[..., 1] ** 2 + U[..., 2] ** 2) / U[..., 0] ** 2))


def functional(mesh: cfd.Mesh, p: cfd.Params, U, faces=None) -> float:
    f = faces if faces is not None else mesh.faces()
    idx = bump_faces(f)
    return float(np.sum(pressure(p, U[f["owner"][idx]]) * f["area"][idx]))


def functional_dU(mesh: cfd.Mesh, p: cfd.Params, U) -> np.ndarray:
    """dJ/dU (N x 6): A dp/dU at each bump face's owner."""
    f = mesh.faces()
    idx = bump_faces(f)
    g = np.zeros_like(U)
    gm1 = p.gamma - 1.0
    for i in idx:
        o = f["owner"][i]
        r, ru, rv = U[o, 0], U[o, 1], U[o, 2]
        u, v = ru / r, rv / r
        A = f["area"][i]
        g[o, 0] += A * gm1 * 0.5 * (u * u + v * v)
        g[o, 1] += A * (-gm1 * u)
        g[o, 2] += A * (-gm1 * v)
        g[o, 3] += A * gm1
    return g


def design_modes(mesh: cfd.Mesh, n_design: int = 64) -> np.ndarray:
    """(n_design, n_nodes + 1) vertical node displacement per unit design variable."""
    x, y, _, _, _ = mesh.desc()
    h = 0.05
    R = (0.25 + h * h) / (2 * h)
    yb = np.where((x > 1.0) & (x < 2.0), np.sqrt(np.maximum(R * R - (x - 1.5) ** 2, 0.0)) - (R - h), 0.0)
    eta = np.clip((y - yb) / (1.0 - yb), 0.0, 1.0)
    decay = 1.0 - eta
    w = 1.0 / (n_design + 1)
    xc = 1.0 + w * (np.arange(n_design) + 1)
    modes = np.maximum(0.0, 1.0 - np.abs(x[None, :] - xc[:, None]) / w) * decay[None, :]
    modes[:, 0] = 0.0
    return modes


def adjoint_solve(solver: cfd.Solver, U, rhs, tol=1e-10, m=60, restarts=40, precond=2):
    """psi with (dR/dU)^T psi = rhs at the steady state (pseudo-time term removed)."""
    solver.compute_residuals(U, True)
    solver.compute_local_time_step(U, 1e30)
    R = np.zeros_like(U)
    psi, it = solver.solve_gmres(U, R, rhs, solver.gmres_opts(m=m, tol=tol, jv=2, precision=64,
                                                             precond=precond, restarts=restarts))
    return psi, it


def design_gradient(mesh: cfd.Mesh, solver: cfd.Solver, p: cfd.Params, U, modes, step=1e-7, tol=1e-10):
    """Adjoint gradient dJ/dalpha for every row of `modes` at the converged state U."""
    psi, it = adjoint_solve(solver, U, functional_dU(mesh, p, U), tol=tol)
    x0, y0, _, _, _ = mesh.desc()
    grad = np.zeros(len(modes))
    try:
        for k, mode in enumerate(modes):
            vals = []
            for sgn in (1.0, -1.0):
                mesh.set_nodes(x0, y0 + sgn * step * mode)
                solver.update_geometry(mesh)
                J = functional(mesh, p, U)
                R = solver.compute_residuals(U, False)
                vals.append((J, R))
            dJ = (vals[0][0] - vals[1][0]) / (2 * step)
            dR = (vals[0][1] - vals[1][1]) / (2 * step)
            grad[k] = dJ - float(np.sum(psi * dR))
    finally:
        mesh.set_nodes(x0, y0)
        solver.update_geometry(mesh)
    return grad, psi, it


def converge(solver: cfd.Solver, U0, rel_drop=1e-12, max_steps=400, jv=1, precision=64, tolerance=0.0,
             cfl0=1.0):
    """Steady state by JFNK with the SER controller and the block SGS preconditioner.

    Stops at a drop of `rel_drop` of the volume-weighted residual norm or below
    the absolute `tolerance` (a re-solve from a nearby converged state needs
    the absolute test: its own initial residual is already small)."""
    g = solver.gmres_opts(m=60, tol=1e-4, jv=jv, precision=precision, precond=2, restarts=2)
    o = solver.newton_opts(gmres=g, tolerance=tolerance, rel_drop=rel_drop, max_steps=max_steps, controller=1,
                           cfl0=cfl0, cfl_max=1e12)
    return solver.solve_implicit(U0, o)
