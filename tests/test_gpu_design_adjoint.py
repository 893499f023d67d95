"""Discrete-adjoint design gradient of the bump wall-pressure integral
(configuration 4) against central finite differences of re-converged flows,
on the configuration-4 mesh at quarter scale (51,639 cells).

The adjoint gives all 64 derivatives (64 bump-node displacement modes) from
one transposed solve.  They are checked, with central differences at the
configuration's step 1e-6:
  * along a seeded random +-1 combination of all 64 modes (the directional
    derivative is the sum of the 64 adjoint derivatives with those signs),
    so every mode enters the check;
  * for two single modes, near each end of the bump.

Bar: 1e-4 relative.  The base flow is converged to a 1e-13 drop and each
perturbed flow to twice the base flow's absolute residual level, so the
convergence error of J divided by 2h stays far below 1e-4 of the
derivative; the O(h^2) truncation is far smaller still.

The flexible-GMRES adjoint driver that the gradient uses is also checked
against the defect-correction driver on a smaller mesh.
"""
import numpy as np
import pytest

from mycode_amd import adjoint, cfd, configs

pytestmark = pytest.mark.gpu

SCALE, H = 0.25, 1e-6


def test_design_gradient_all_modes_vs_central_fd():
    mesh, p, U0 = configs.setup(4, SCALE)
    s = cfd.Solver(mesh, p)
    U, res, hist = adjoint.converge(s, U0, rel_drop=1e-13)
    print(f"N={mesh.n_cells}: converged={res.converged} steps={res.steps} drop={hist['vm'][-1] / hist['vm'][0]:.2e}",
          flush=True)
    assert res.converged
    modes = adjoint.design_modes(mesh, 64)
    g_adj, psi, it = adjoint.design_gradient(mesh, s, p, U, modes)
    print("adjoint GMRES iterations", it, flush=True)
    assert g_adj.shape == (64,) and np.all(np.isfinite(g_adj))
    d = np.random.default_rng(64).choice([-1.0, 1.0], 64)
    checks = [("all 64 modes, random signs", d @ modes, float(d @ g_adj)), ("mode 3", modes[3], g_adj[3]),
              ("mode 60", modes[60], g_adj[60])]
    x0, y0, _, _, _ = mesh.desc()
    tol = 2.0 * res.res_vm
    for name, mode, ga in checks:
        J = []
        for sgn in (1.0, -1.0):
            m2 = cfd.Mesh.generate(4, SCALE)
            m2.set_nodes(x0, y0 + sgn * H * mode)
            s2 = cfd.Solver(m2, p)
            U2, r2, _ = adjoint.converge(s2, U, rel_drop=0.0, tolerance=tol, cfl0=100.0, newton_below=1.0)
            assert r2.converged
            J.append(adjoint.functional(m2, p, U2))
        fd = (J[0] - J[1]) / (2 * H)
        print(f"{name}: adjoint {ga:.10e} FD {fd:.10e} rel {abs(ga - fd) / abs(fd):.2e}", flush=True)
        assert abs(ga - fd) <= 1e-4 * abs(fd)


def test_fgmres_adjoint_matches_defect_correction():
    """The two adjoint drivers on the configuration-4 mesh at 0.1 scale: the
    flexible GMRES (outer FP64 Krylov space over inner transposed solves)
    and the pseudo-time defect correction each reach a relative residual of
    1e-10, recomputed here with the steady reverse product through the host
    entry point, and their solutions agree to 1e-8 relative (the two
    residual bars allow about the condition number times 1e-10)."""
    mesh, p, U0 = configs.setup(4, 0.1)
    s = cfd.Solver(mesh, p)
    U, res, _ = adjoint.converge(s, U0, rel_drop=1e-12)
    assert res.converged
    rhs = adjoint.functional_dU(mesh, p, U)
    psi_f, inner, outer, rel_f = adjoint.adjoint_solve_fgmres(s, U, rhs)
    psi_d, its, k, rel_d = adjoint.adjoint_solve(s, U, rhs)
    print(f"N={mesh.n_cells}: fgmres {outer} outer / {inner} inner ({rel_f:.2e}); "
          f"defect correction {k} outer / {its} inner ({rel_d:.2e})", flush=True)
    s.compute_residuals(U, True)
    s.compute_local_time_step(U, 1e30)  # the time term vanishes
    bn = np.linalg.norm(rhs)
    for name, psi in (("fgmres", psi_f), ("defect correction", psi_d)):
        r = np.linalg.norm(s.vjp_adjoint(U, psi)[:, :p.nvar] - rhs[:, :p.nvar]) / bn
        print(f"{name}: true relative residual {r:.2e}", flush=True)
        assert r <= 1.5e-10
    assert np.linalg.norm(psi_f - psi_d) <= 1e-8 * np.linalg.norm(psi_d)
    assert inner < its


def test_newton_fgmres_tangent_and_fd_products():
    """The Newton stage of adjoint.converge on the configuration-4 mesh at 0.1
    scale, from the pseudo-time iteration's state at a 1e-4 drop: with the
    tangent product (FP32 inner Krylov) and with the finite-difference product
    (FP64 inner Krylov) the norm falls by at least 10x per step (the first
    step, from the pseudo-time state, about 60x; the later ones by the linear
    tolerance, 1e-3), both reach a 1e-11 drop in at most four steps, and the
    two converged states agree to 1e-9 relative."""
    mesh, p, U0 = configs.setup(4, 0.1)
    s = cfd.Solver(mesh, p)
    U1, res, hist = adjoint.converge(s, U0, rel_drop=1e-4, newton_below=0.0)
    assert res.converged
    target = 1e-11 * float(hist["vm"][0])
    out = {}
    for jv, prec in ((1, 32), (0, 64)):
        U, k, inner, vms, ok = adjoint.newton_fgmres(s, U1, target, jv=jv, precision=prec)
        print(f"jv={jv}: {k} Newton steps, {inner} inner iterations, norms {['%.2e' % v for v in vms]}", flush=True)
        assert ok and vms[-1] <= target and k <= 4
        norms = [res.res_vm] + vms
        assert all(b <= 0.1 * a for a, b in zip(norms[:-1], norms[1:]))
        out[jv] = U
    assert np.linalg.norm(out[0] - out[1]) <= 1e-9 * np.linalg.norm(out[1])
