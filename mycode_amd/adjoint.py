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


def _fgmres(apply_A, apply_M, b, x, tol, V, Z, w, max_outer, log=None, log_every=25, what="fgmres"):
    """Flexible GMRES on device vectors (torch FP64, the library's layout):
    x += the correction that brings |b - A x| to tol |b| or below, restarted
    every Z.shape[0] steps, at most max_outer steps in all.  Step j
    preconditions the basis vector v_j by z_j = apply_M(v_j) (any
    approximate solve; apply_M(v, z) writes z and returns its own inner
    iteration count) and orthogonalises A z_j by classical Gram-Schmidt with
    one reorthogonalisation; the Hessenberg least squares (Givens rotations)
    runs on the host in FP64.  V ((m + 1) x n), Z (m x n) and w (n) are
    the caller's work space.  Returns (inner iterations, outer steps,
    relative residual |b - A x| / |b| recomputed at the end)."""
    import torch

    m = Z.shape[0]
    bn = float(torch.linalg.norm(b))
    if bn == 0.0:
        return 0, 0, 0.0
    if bool(torch.any(x != 0)):
        apply_A(x, w)
        r = b - w
    else:  # no product of a zero vector: the finite-difference product scales its step by 1 / |v|
        r = b.clone()
    rn = float(torch.linalg.norm(r))
    inner, outer = 0, 0
    while rn > tol * bn and outer < max_outer:
        V[0] = r / rn
        H = np.zeros((m + 1, m))
        g = np.zeros(m + 1)
        g[0] = rn
        cs, sn = np.zeros(m), np.zeros(m)
        j_end = 0
        for j in range(m):
            Z[j].zero_()
            inner += apply_M(V[j], Z[j])
            apply_A(Z[j], w)
            h = torch.mv(V[: j + 1], w)  # classical Gram-Schmidt, twice
            w -= torch.mv(V[: j + 1].t(), h)
            h2 = torch.mv(V[: j + 1], w)
            w -= torch.mv(V[: j + 1].t(), h2)
            hn = float(torch.linalg.norm(w))
            H[: j + 1, j] = (h + h2).cpu().numpy()
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
                log(f"{what} {outer}: estimate {abs(g[j + 1]) / bn:.3e}, inner {inner}")
            if abs(g[j + 1]) <= tol * bn or hn == 0.0 or outer >= max_outer:
                break
        y = np.linalg.solve(np.triu(H[:j_end, :j_end]), g[:j_end])
        x += torch.mv(Z[:j_end].t(), torch.from_numpy(y).to(x.device))
        apply_A(x, w)
        r = b - w
        rn = float(torch.linalg.norm(r))
    return inner, outer, rn / bn


def _to_device(solver: cfd.Solver, a6):
    """(N x 6) host array -> device vector in the library's layout."""
    import torch

    t6 = torch.from_numpy(np.ascontiguousarray(a6, np.float64)).to("cuda")
    out = torch.zeros(solver.params().nvar * solver.pitch, dtype=torch.float64, device="cuda")
    solver.dev_pack(t6.data_ptr(), out.data_ptr())
    torch.cuda.current_stream().synchronize()  # t6 may be freed on return
    return out


def _to_host(solver: cfd.Solver, x):
    import torch

    x6 = torch.zeros((solver.N, 6), dtype=torch.float64, device="cuda")
    solver.dev_unpack(x.data_ptr(), x6.data_ptr())
    return x6.cpu().numpy()


def adjoint_solve_fgmres(solver: cfd.Solver, U, rhs, tol=1e-10, cfl=1e6, m_outer=200, max_outer=1000, m=60,
                         restarts=0, inner_tol=1e-2, precond=2, precision=32, log=None, log_every=25):
    """psi with (dR/dU)^T psi = rhs at the steady state U by flexible GMRES
    (FGMRES(m_outer), FP64 basis): each outer step's preconditioner is one
    approximate solve of the shifted transposed system
    (dR/dU^T + V/dtau(cfl)) z = v by the library's GMRES (``jv = 2``,
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

    n = solver.params().nvar * solver.pitch
    solver.set_stream(torch.cuda.current_stream().cuda_stream)
    try:
        Ud = _to_device(solver, U)
        rhs_d = _to_device(solver, rhs)
        zero = torch.zeros_like(Ud)
        solver.dev_residual(Ud.data_ptr(), zero.data_ptr(), True)  # forms the limiter the products hold frozen
        zero.zero_()
        opts = solver.gmres_opts(m=m, tol=inner_tol, jv=2, precision=precision, precond=precond, restarts=restarts)

        def apply_A(x, out):  # the steady reverse product: the time term vanishes
            solver.dev_local_timestep(Ud.data_ptr(), 1e30)
            solver.dev_vjp_adjoint(Ud.data_ptr(), x.data_ptr(), out.data_ptr())

        def apply_M(v, z):
            solver.dev_local_timestep(Ud.data_ptr(), cfl)
            return solver.dev_gmres(Ud.data_ptr(), zero.data_ptr(), v.data_ptr(), z.data_ptr(), opts)

        V = torch.zeros((m_outer + 1, n), dtype=torch.float64, device="cuda")
        Z = torch.zeros((m_outer, n), dtype=torch.float64, device="cuda")
        w = torch.zeros(n, dtype=torch.float64, device="cuda")
        psi = torch.zeros_like(Ud)
        inner, outer, rel = _fgmres(apply_A, apply_M, rhs_d, psi, tol, V, Z, w, max_outer, log, log_every,
                                    "adjoint fgmres")
        out = _to_host(solver, psi)
    finally:
        solver.set_stream(None)
    return out, inner, outer, rel


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


def newton_fgmres(solver: cfd.Solver, U, vm_target, jv=1, precision=32, precond=2, cfl=1e6, eta=1e-3,
                  m_outer=100, max_outer=600, max_newton=12, log=None):
    """Newton steps on R(U) = 0 from U (N x 6) to a volume-weighted residual
    norm of vm_target.  Each step solves dR/dU dU = -R by FGMRES(m_outer) in
    FP64 (the outer product: the tangent (jv 1) or finite-difference (jv 0)
    Jacobian-vector product with the pseudo-time term switched off), each
    outer step preconditioned by one GMRES(60) solve of the shifted system at
    CFL `cfl` (`precond`, `precision` basis), to a relative tolerance of
    min(0.1, max(eta, 0.3 vm_target / vm)), so that the last step is solved
    no tighter than the target needs; then a backtracking search over the step lengths
    1, 1/2, ..., 1/16 keeps the first that lowers the norm.

    Near a converged state the Newton steps shed the pseudo-transient
    iteration's many capped GMRES solves: on the configuration-4 bump at full
    size, from a 1e-4 drop, three steps of about 5,000 inner iterations each
    reached a 1e-12 drop in 9 s, where the pseudo-time iteration took 324
    more steps of up to 180 GMRES iterations (33 s).  Returns (U, Newton
    steps, inner GMRES iterations, the norms after each step, ok); ok is
    False when a step
    found no length that lowers the norm (U is then the last accepted state)
    or lowered it by less than half."""
    import torch

    nv, n = solver.params().nvar, solver.params().nvar * solver.pitch
    solver.set_stream(torch.cuda.current_stream().cuda_stream)
    try:
        Ud = _to_device(solver, U)
        Rd = torch.zeros_like(Ud)
        out = torch.zeros(1, dtype=torch.float64, device="cuda")
        opts = solver.gmres_opts(m=60, tol=1e-2, jv=jv, precision=precision, precond=precond, restarts=0)

        def residual():
            solver.dev_residual(Ud.data_ptr(), Rd.data_ptr(), True)
            solver.dev_norm_vw_noalloc(Rd.data_ptr(), out.data_ptr())
            return float(out.item())

        def apply_A(x, y):
            solver.dev_local_timestep(Ud.data_ptr(), 1e30)  # the time term vanishes
            if jv == 0:
                solver.dev_jv_fd(Ud.data_ptr(), Rd.data_ptr(), x.data_ptr(), y.data_ptr())
            else:
                solver.dev_jv_tangent(Ud.data_ptr(), x.data_ptr(), y.data_ptr(), 64)

        def apply_M(v, z):
            solver.dev_local_timestep(Ud.data_ptr(), cfl)
            return solver.dev_gmres(Ud.data_ptr(), Rd.data_ptr(), v.data_ptr(), z.data_ptr(), opts)

        V = torch.zeros((m_outer + 1, n), dtype=torch.float64, device="cuda")
        Z = torch.zeros((m_outer, n), dtype=torch.float64, device="cuda")
        w = torch.zeros(n, dtype=torch.float64, device="cuda")
        vm = residual()
        steps, inner, ok, vms = 0, 0, True, []
        while vm > vm_target and steps < max_newton:
            tol = min(0.1, max(eta, 0.3 * vm_target / vm))
            dU = torch.zeros_like(Ud)
            it, outer, rel = _fgmres(apply_A, apply_M, -Rd, dU, tol, V, Z, w, max_outer)
            inner += it
            U_old = Ud.clone()
            lam, vm_new = 1.0, float("inf")
            while lam >= 1.0 / 16:
                torch.add(U_old, dU, alpha=lam, out=Ud)
                vm_new = residual()
                if vm_new < vm:
                    break
                lam *= 0.5
            steps += 1
            if log:
                log(f"newton {steps}: vm {vm:.3e} -> {vm_new:.3e}, linear {rel:.1e} in {outer} outer / {it} inner, "
                    f"step length {lam:g}")
            if not vm_new < vm:
                Ud.copy_(U_old)
                residual()
                ok = False
                break
            ok = vm_new <= 0.5 * vm
            vm = vm_new
            vms.append(vm)
            if not ok:
                break
        U_out = _to_host(solver, Ud)
    finally:
        solver.set_stream(None)
    out_U = np.array(U, dtype=np.float64, copy=True)
    out_U[:, :nv] = U_out[:, :nv]
    return out_U, steps, inner, vms, ok


def converge(solver: cfd.Solver, U0, rel_drop=1e-10, max_steps=600, jv=1, precision=32, tolerance=0.0,
             cfl0=1.0, log_every=0, newton_below=1e-4):
    """Steady state by JFNK: the tangent product (FP32 matvec, preconditioner
    and basis by default), block SGS, GMRES(60) with two restarts at a 1e-2
    tolerance and the adaptive CFL controller, then (newton_below > 0)
    Newton steps solved by FGMRES (newton_fgmres) from a drop of
    newton_below on; newton_below = 1 goes straight to them (a re-solve from
    a nearby converged state).  Should a Newton step fail to halve the norm,
    or the Newton steps end short of the target, the pseudo-time iteration
    resumes from the last accepted state.

    Stops at a drop of `rel_drop` of the volume-weighted residual norm or below
    the absolute `tolerance` (a re-solve from a nearby converged state needs
    the absolute test: its own initial residual is already small).
    log_every > 1 prints the progress to stderr every that many steps.
    Returns (U, result, history) as Solver.solve_implicit does, with the
    Newton steps counted in result.steps, result.total_gmres_iters (their
    inner iterations) and result.seconds, and appended to history['vm']."""
    import sys
    import time

    g = solver.gmres_opts(m=60, tol=1e-2, jv=jv, precision=precision, precond=2, restarts=2)

    def pseudo_time(U, drop, steps, cfl, tol):
        o = solver.newton_opts(gmres=g, tolerance=tol, rel_drop=drop, max_steps=steps, controller=2,
                               cfl0=cfl, cfl_max=1e12, log=log_every)
        return solver.solve_implicit(U, o)

    if newton_below <= 0.0:
        return pseudo_time(U0, rel_drop, max_steps, cfl0, tolerance)
    U, res, hist = pseudo_time(U0, max(rel_drop, newton_below), max_steps, cfl0, tolerance)
    vm0 = float(hist["vm"][0])
    target = max(rel_drop * vm0, tolerance)
    if res.res_vm <= target or res.steps >= max_steps:
        return U, res, hist
    log = (lambda m: print("newton-fgmres: " + m, file=sys.stderr, flush=True)) if log_every > 1 else None
    t0 = time.perf_counter()
    U, k, inner, vms, ok = newton_fgmres(solver, U, target, jv=jv, precision=precision, log=log)
    res.seconds += time.perf_counter() - t0
    res.steps += k
    res.total_gmres_iters += inner
    if vms:
        res.res_vm = vms[-1]
    vm = res.res_vm
    hist = dict(hist)
    hist["vm"] = np.concatenate([hist["vm"], vms])
    if vm <= target:
        res.converged = 1
        return U, res, hist
    if res.steps >= max_steps:
        res.converged = 0
        return U, res, hist
    cfl = float(hist["cfl"][-1]) if len(hist["cfl"]) else cfl0
    U, r2, h2 = pseudo_time(U, 0.0, max_steps - res.steps, cfl, target)  # from the last accepted Newton state
    r2.steps += res.steps
    r2.total_gmres_iters += res.total_gmres_iters
    r2.seconds += res.seconds
    r2.res_vm0 = res.res_vm0
    conv = r2.res_vm <= target
    r2.converged = int(conv)
    h2 = dict(h2)
    h2["vm"] = np.concatenate([hist["vm"], h2["vm"]])
    return U, r2, h2
