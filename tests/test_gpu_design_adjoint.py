"""Discrete-adjoint design gradient of the bump wall-pressure integral
(configuration 4) against central finite differences of re-converged flows
on a coarse mesh.

Bar: 1e-4 relative per design variable.  The FD step is 1e-5: the
perturbed flows are re-converged to the base flow's absolute residual
level, whose error in J divided by 2h stays below 1e-5 of the gradient,
and the O(h^2) truncation is far smaller.
"""
import numpy as np
import pytest

from mycode_amd import adjoint, cfd, configs

pytestmark = pytest.mark.gpu


def test_design_gradient_vs_central_fd():
    mesh, p, U0 = configs.setup(4, 0.05)
    s = cfd.Solver(mesh, p)
    U, res, hist = adjoint.converge(s, U0)
    print(f"N={mesh.n_cells}: converged={res.converged} steps={res.steps} drop={hist['vm'][-1] / hist['vm'][0]:.2e}")
    assert res.converged
    modes = adjoint.design_modes(mesh, 64)
    pick = [5, 20, 31, 47]
    g_adj, psi, it = adjoint.design_gradient(mesh, s, p, U, modes[pick])
    print("adjoint GMRES iterations", it, "gradient", g_adj)
    x0, y0, _, _, _ = mesh.desc()
    h = 1e-5
    tol = 2.0 * res.res_vm
    for k, row in zip(pick, g_adj):
        J = []
        for sgn in (1.0, -1.0):
            m2 = cfd.Mesh.generate(4, 0.05)
            m2.set_nodes(x0, y0 + sgn * h * modes[k])
            s2 = cfd.Solver(m2, p)
            U2, r2, _ = adjoint.converge(s2, U, rel_drop=0.0, tolerance=tol, cfl0=100.0)
            assert r2.converged
            J.append(adjoint.functional(m2, p, U2))
        fd = (J[0] - J[1]) / (2 * h)
        print(f"design {k}: adjoint {row:.10e} FD {fd:.10e}")
        assert abs(row - fd) <= 1e-4 * max(abs(fd), 1e-3 * np.abs(g_adj).max())
