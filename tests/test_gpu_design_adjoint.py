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
            U2, r2, _ = adjoint.converge(s2, U, rel_drop=0.0, tolerance=tol, cfl0=100.0)
            assert r2.converged
            J.append(adjoint.functional(m2, p, U2))
        fd = (J[0] - J[1]) / (2 * H)
        print(f"{name}: adjoint {ga:.10e} FD {fd:.10e} rel {abs(ga - fd) / abs(fd):.2e}", flush=True)
        assert abs(ga - fd) <= 1e-4 * abs(fd)
