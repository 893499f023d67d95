"""Discrete-adjoint design gradients on the GPU (no reference counterpart).

For a converged state U*(alpha) with R(U*, X(alpha)) = 0 and a functional
J(U, X), the total derivative is

    dJ/dalpha_k = dJ/dX . dX/dalpha_k  -  psi^T (dR/dX . dX/dalpha_k),
    (dR/dU)^T psi = (dJ/dU)^T,

psi from flexible GMRES on the transposed system, preconditioned by GMRES
on the shifted transposed system (the library's reverse product, ``jv = 2``), and the two partials at fixed U by central differences of the
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


def adjoint_solve(solver: cfd.Solver, U, rhs, tol=1e-10, cfl0=10.0, cfl_max=1e6, m=60, restarts=2,
                  inner_tol=1e-2, max_outer=600, precond=2, precision=32, log=None, log_every=25):
    """psi with (dR/dU)^T psi = rhs at the steady state U, by defect correction
    in pseudo-time, the adjoint counterpart of the Newton iteration that found
    U:  psi <- psi + (dR/dU^T + V/dtau)^-1 (rhs - dR/dU^T psi), the shifted
    transposed system solved by GMRES (the library's reverse product,
    ``jv = 2``) with the block SGS preconditioner; the inner solve's basis and
    preconditioner run in `precision` (the reverse product itself in FP64),
    the defect and psi in FP64.  Every vector stays on the device (torch
    tensors in the library's layout, on torch's stream).  The FP32 basis
    suits systems like the subsonic bump's; where the shifted Jacobian's
    condition number is far above 1/eps32 (the supersonic wedge's is near
    6.5e12 at CFL 20) the inner solve needs ``precision=64``.

    The CFL follows the rule of the library's adaptive controller: x2 after
    an iteration that lowered the defect by 1 % or more with GMRES at its
    tolerance (x1.25 when GMRES missed it), x1.1 (x0.7) after a smaller change
    up to a 20 % rise, x0.5 after a larger rise; a rise above 10x is undone
    and the CFL cut tenfold; the CFL stays in [1e-2 cfl0, cfl_max].  Stops at
    |defect| <= tol |rhs|.  Returns (psi (N x 6, reference order), total
    GMRES iterations, outer iterations, relative defect)."""
    import torch

    dev = torch.device("cuda")
    nv, P, N = solver.params().nvar, solver.pitch, solver.N
    solver.set_stream(torch.cuda.current_stream().cuda_stream)
    try:
        def to_dev(a6):
            t6 = torch.from_numpy(np.ascontiguousarray(a6, np.float64)).to(dev)
            out = torch.zeros(nv * P, dtype=torch.float64, device=dev)
            solver.dev_pack(t6.data_ptr(), out.data_ptr())
            return out, t6

        Ud, _k1 = to_dev(U)
        rhs_d, _k2 = to_dev(rhs)
        zero = torch.zeros_like(Ud)
        solver.dev_residual(Ud.data_ptr(), zero.data_ptr(), True)  # forms the limiter the products hold frozen
        zero.zero_()
        psi = torch.zeros_like(Ud)
        r = rhs_d.clone()
        x = torch.zeros_like(Ud)
        Ax = torch.zeros_like(Ud)
        bn = float(torch.linalg.norm(rhs_d))
        rn = bn
        cfl, its, k = cfl0, 0, 0
        cap = m * (restarts + 1)
        opts = solver.gmres_opts(m=m, tol=inner_tol, jv=2, precision=precision, precond=precond, restarts=restarts)
        for k in range(max_outer):
            if rn <= tol * bn:
                break
            solver.dev_local_timestep(Ud.data_ptr(), cfl)
            it = solver.dev_gmres(Ud.data_ptr(), zero.data_ptr(), r.data_ptr(), x.data_ptr(), opts)
            its += it
            gmres_ok = it < cap
            trial = psi + x
            solver.dev_local_timestep(Ud.data_ptr(), 1e30)  # steady operator: the time term vanishes
            solver.dev_vjp_adjoint(Ud.data_ptr(), trial.data_ptr(), Ax.data_ptr())
            r_new = rhs_d - Ax
            rn_new = float(torch.linalg.norm(r_new))
            if not np.isfinite(rn_new) or rn_new > 10.0 * rn:
                cfl = max(0.1 * cfl, 1e-2 * cfl0)
                continue
            if rn_new <= 0.99 * rn:
                cfl *= 2.0 if gmres_ok else 1.25
            elif rn_new <= 1.2 * rn:
                cfl *= 1.1 if gmres_ok else 0.7
            else:
                cfl *= 0.5
            cfl = min(max(cfl, 1e-2 * cfl0), cfl_max)
            psi, r, rn = trial, r_new, rn_new
            if log and k % log_every == 0:
                log(f"adjoint outer {k}: defect {rn / bn:.3e}, cfl {cfl:.3g}, gmres {it}")
        psi6 = torch.zeros((N, 6), dtype=torch.float64, device=dev)
        solver.dev_unpack(psi.data_ptr(), psi6.data_ptr())
        out = psi6.cpu().numpy()
    finally:
        solver.set_stream(None)
    return out, its, k, rn / bn


def adjoint_solve_fgmres(solver: cfd.Solver, U, rhs, tol=1e-10, cfl=1e6, m_outer=200, max_outer=1000, m=60,
                         restarts=0, inner_tol=1e-2, precond=2, precision=32, log=None, log_every=25):
    """psi with (dR/dU)^T psi = rhs at the steady state U by flexible GMRES
    (FGMRES(m_outer), FP64 basis, classical Gram-Schmidt with one
    reorthogonalisation, Hessenberg least squares on the host): each outer
    step's preconditioner is one approximate solve of the shifted transposed
    system (dR/dU^T + V/dtau(cfl)) z = v by the library's GMRES (``jv = 2``,
    GMRES(m) with `restarts` at `inner_tol`, block SGS, `precision` basis), so
    the outer Krylov space accelerates the pseudo-time iteration that
    adjoint_solve runs as plain defect correction.  The outer products are
    the steady reverse product.  Every vector stays on the device; the outer
    space holds 2 m_outer + 1 FP64 vectors (10.6 GB for the 824,378-cell bump
    at m_outer = 200), sized so that the outer iteration rarely restarts.

    On the configuration-4 bump at full size (tol 1e-10) the defaults took
    170 outer steps and 10,200 inner iterations, where the defect correction
    needed 376 outer steps and 67,041 inner iterations; restarting the outer
    iteration every 60 or 100 steps cost 2.2x or 1.6x the inner iterations.
    Returns (psi (N x 6, reference order), inner GMRES iterations, outer
    iterations, relative residual |rhs - dR/dU^T psi| / |rhs|)."""
    import torch

    dev = torch.device("cuda")
    nv, P, N = solver.params().nvar, solver.pitch, solver.N
    n = nv * P
    solver.set_stream(torch.cuda.current_stream().cuda_stream)
    try:
        def to_dev(a6):
            t6 = torch.from_numpy(np.ascontiguousarray(a6, np.float64)).to(dev)
            out = torch.zeros(n, dtype=torch.float64, device=dev)
            solver.dev_pack(t6.data_ptr(), out.data_ptr())
            return out, t6

        Ud, _k1 = to_dev(U)
        rhs_d, _k2 = to_dev(rhs)
        zero = torch.zeros_like(Ud)
        solver.dev_residual(Ud.data_ptr(), zero.data_ptr(), True)  # forms the limiter the products hold frozen
        zero.zero_()
        opts = solver.gmres_opts(m=m, tol=inner_tol, jv=2, precision=precision, precond=precond, restarts=restarts)
        V = torch.zeros((m_outer + 1, n), dtype=torch.float64, device=dev)
        Z = torch.zeros((m_outer, n), dtype=torch.float64, device=dev)
        w = torch.zeros(n, dtype=torch.float64, device=dev)
        psi = torch.zeros_like(Ud)
        bn = float(torch.linalg.norm(rhs_d))
        r = rhs_d.clone()
        rn = bn
        inner, outer = 0, 0

        def apply_steady(x, out):
            solver.dev_local_timestep(Ud.data_ptr(), 1e30)  # the time term vanishes
            solver.dev_vjp_adjoint(Ud.data_ptr(), x.data_ptr(), out.data_ptr())

        while rn > tol * bn and outer < max_outer:
            V[0] = r / rn
            H = np.zeros((m_outer + 1, m_outer))
            g = np.zeros(m_outer + 1)
            g[0] = rn
            cs, sn = np.zeros(m_outer), np.zeros(m_outer)
            j_end = 0
            for j in range(m_outer):
                solver.dev_local_timestep(Ud.data_ptr(), cfl)
                Z[j].zero_()
                inner += solver.dev_gmres(Ud.data_ptr(), zero.data_ptr(), V[j].data_ptr(), Z[j].data_ptr(), opts)
                apply_steady(Z[j], w)
                h = torch.mv(V[: j + 1], w)  # classical Gram-Schmidt, twice
                w -= torch.mv(V[: j + 1].t(), h)
                h2 = torch.mv(V[: j + 1], w)
                w -= torch.mv(V[: j + 1].t(), h2)
                hc = (h + h2).cpu().numpy()
                hn = float(torch.linalg.norm(w))
                H[: j + 1, j] = hc
                H[j + 1, j] = hn
                if hn > 0.0:
                    V[j + 1] = w / hn
                for i in range(j):  # the earlier rotations on the new column
                    t = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
                    H[i + 1, j] = -sn[i] * H[i, j] + cs[i] * H[i + 1, j]
                    H[i, j] = t
                den = np.hypot(H[j, j], H[j + 1, j])
                cs[j], sn[j] = (1.0, 0.0) if den == 0.0 else (H[j, j] / den, H[j + 1, j] / den)
                H[j, j] = den
                H[j + 1, j] = 0.0
                g[j + 1] = -sn[j] * g[j]
                g[j] = cs[j] * g[j]
                outer += 1
                j_end = j + 1
                if log and outer % log_every == 0:
                    log(f"adjoint fgmres {outer}: estimate {abs(g[j + 1]) / bn:.3e}, inner {inner}")
                if abs(g[j + 1]) <= tol * bn or hn == 0.0 or outer >= max_outer:
                    break
            y = np.linalg.solve(np.triu(H[:j_end, :j_end]), g[:j_end]) if j_end else np.zeros(0)
            psi += torch.mv(Z[:j_end].t(), torch.from_numpy(y).to(dev))
            apply_steady(psi, w)
            r = rhs_d - w
            rn = float(torch.linalg.norm(r))
        psi6 = torch.zeros((N, 6), dtype=torch.float64, device=dev)
        solver.dev_unpack(psi.data_ptr(), psi6.data_ptr())
        out = psi6.cpu().numpy()
    finally:
        solver.set_stream(None)
    return out, inner, outer, rn / bn


def gradient_from_adjoint(mesh: cfd.Mesh, solver: cfd.Solver, p: cfd.Params, U, modes, psi, step=1e-7):
    """dJ/dalpha_k = dJ/dX . dX/dalpha_k - psi^T (dR/dX . dX/dalpha_k) for every
    row of `modes`, the two partials at fixed U by central differences of the
    node coordinates with `step`."""
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
    return grad


def design_gradient(mesh: cfd.Mesh, solver: cfd.Solver, p: cfd.Params, U, modes, step=1e-7, tol=1e-10, log=None):
    """Adjoint gradient dJ/dalpha for every row of `modes` at the converged
    state U; returns (gradient, psi, adjoint GMRES iterations)."""
    psi, it, _, _ = adjoint_solve_fgmres(solver, U, functional_dU(mesh, p, U), tol=tol, log=log)
    return gradient_from_adjoint(mesh, solver, p, U, modes, psi, step), psi, it


def converge(solver: cfd.Solver, U0, rel_drop=1e-10, max_steps=600, jv=1, precision=32, tolerance=0.0,
             cfl0=1.0, log_every=0):
    """Steady state by JFNK: the tangent product (FP32 matvec, preconditioner
    and basis by default), block SGS, GMRES(60) with two restarts at a 1e-2
    tolerance and the adaptive CFL controller.

    Stops at a drop of `rel_drop` of the volume-weighted residual norm or below
    the absolute `tolerance` (a re-solve from a nearby converged state needs
    the absolute test: its own initial residual is already small).
    log_every > 1 prints the Newton progress to stderr every that many steps."""
    g = solver.gmres_opts(m=60, tol=1e-2, jv=jv, precision=precision, precond=2, restarts=2)
    o = solver.newton_opts(gmres=g, tolerance=tolerance, rel_drop=rel_drop, max_steps=max_steps, controller=2,
                           cfl0=cfl0, cfl_max=1e12, log=log_every)
    return solver.solve_implicit(U0, o)
