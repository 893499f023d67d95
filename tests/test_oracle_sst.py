"""SST k-omega in the oracle (no reference counterpart; the reference only
reserves the rhok / rhow slots and an uncalled wall-distance routine).

Pinned by: the wall distance against a brute-force numpy evaluation of the
reference's definition (Solver.cpp:1981-2015); the source terms on a uniform
state against the published SST-2003 closed form evaluated independently in
numpy; freestream preservation of the mean-flow equations; complex step vs
central differences (test_oracle_derivatives.py covers configuration 3).
"""
import math

import numpy as np

from mycode_amd import configs
from oracle import oracle


def _setup(scale=0.03):
    mesh, p, U = configs.setup(3, scale)
    o = oracle.OracleSolver(oracle.OracleMesh.from_mesh(mesh), p)
    return mesh, p, U, o


def test_wall_distance_brute_force():
    mesh, p, U, o = _setup()
    f, c = mesh.faces(), mesh.cells()
    w = (f["neighbor"] < 0) & (f["bc_tag"] == 4)
    d2 = (c["x_c"][:, None] - f["x_mid"][w][None, :]) ** 2 + (c["y_c"][:, None] - f["y_mid"][w][None, :]) ** 2
    assert np.array_equal(o.wall_distance(), np.sqrt(d2.min(1)))


def test_uniform_state_sources_closed_form():
    mesh, p, U, o = _setup()
    R = o.compute_residuals(U)
    f, c = mesh.faces(), mesh.cells()
    # cells whose faces are all interior and whose neighbours' faces are interior too
    bnd = np.zeros(mesh.n_cells, bool)
    b = f["neighbor"] < 0
    bnd[f["owner"][b]] = True
    ring = bnd.copy()
    inter = ~b
    ring[f["owner"][inter][bnd[f["neighbor"][inter]]]] = True
    ring[f["neighbor"][inter][bnd[f["owner"][inter]]]] = True
    cells = np.where(~ring)[0]
    assert len(cells) > 100
    rho, k, w = U[0, 0], U[0, 4] / U[0, 0], U[0, 5] / U[0, 0]
    nu, d = p.nu, o.wall_distance()[cells]
    bs, b1, b2, sw1, sw2, kap = 0.09, 0.075, 0.0828, 0.5, 0.856, 0.41
    CD = 1e-10
    arg1 = np.minimum(np.maximum(math.sqrt(k) / (bs * w * d), 500 * nu / (d * d * w)), 4 * rho * sw2 * k / (CD * d * d))
    F1 = np.tanh(arg1 ** 4)
    beta = F1 * b1 + (1 - F1) * b2
    V = c["volume"][cells]
    Rk = V * bs * rho * w * k          # production vanishes with the strain
    Rw = V * beta * rho * w * w
    flux_scale = np.abs(R[:, 1]).max() + 1.0
    assert np.max(np.abs(R[cells, 4] - Rk) / Rk) < 1e-9
    assert np.max(np.abs(R[cells, 5] - Rw) / Rw) < 1e-9
    # mean-flow freestream preservation on the same cells
    assert np.max(np.abs(R[cells, :4])) < 1e-9 * flux_scale


def _kw_state(p, U, seed, vary_kw):
    rng = np.random.default_rng(seed)
    U = U.copy()
    U[:, :4] *= 1 + 0.01 * rng.standard_normal((U.shape[0], 4))
    if vary_kw:
        U[:, 4] *= 1 + 0.5 * rng.random(U.shape[0])
        U[:, 5] *= 1 + 0.5 * rng.random(U.shape[0])
    else:  # uniform specific k and omega over a varying density
        U[:, 4] = U[:, 0] * (p.turb_k_inf * 3.0)
        U[:, 5] = U[:, 0] * (p.turb_w_inf * 0.5)
    return U


def test_kw_first_order_convection_invariants():
    """turb_first_order = 1 convects k and omega with the upwind cell values
    and changes nothing else: (a) with uniform specific k and omega the
    Green-Gauss gradients of k and omega vanish (to round-off) away from the
    boundaries, so the reconstruction adds nothing and both settings agree
    there; (b) for varying k and omega the mean-flow rows agree bit for bit
    (the reconstructed k and omega only feed the k and omega convective
    fluxes) while the k and omega rows differ."""
    mesh, p, U0 = configs.setup(3, 0.03)
    om = oracle.OracleMesh.from_mesh(mesh)
    res = {}
    for vary in (False, True):
        U = _kw_state(p, U0, 5, vary)
        for tfo in (0, 1):
            p.turb_first_order = tfo
            res[vary, tfo] = oracle.OracleSolver(om, p).compute_residuals(U)
    # cells with no boundary face and no neighbour with one (boundary ghost
    # states give k and omega gradients there)
    f = mesh.faces()
    b = f["neighbor"] < 0
    bnd = np.zeros(mesh.n_cells, bool)
    bnd[f["owner"][b]] = True
    ring = bnd.copy()
    ring[f["owner"][~b][bnd[f["neighbor"][~b]]]] = True
    ring[f["neighbor"][~b][bnd[f["owner"][~b]]]] = True
    cells = np.where(~ring)[0]
    assert len(cells) > 100
    a0, a1 = res[False, 0][cells, 4:6], res[False, 1][cells, 4:6]
    assert np.max(np.abs(a0 - a1) / np.abs(a0).max(0)) < 1e-12
    assert np.array_equal(res[False, 0][:, :4], res[False, 1][:, :4])
    assert np.array_equal(res[True, 0][:, :4], res[True, 1][:, :4])
    d = np.abs(res[True, 0][:, 4:6] - res[True, 1][:, 4:6])
    assert (d > 0).mean(0).min() > 0.5
