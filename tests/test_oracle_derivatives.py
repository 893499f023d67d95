"""Pins the oracle's complex-step derivative, which in turn pins the GPU
tangent and adjoint kernels.

The complex-step directional derivative Im R(U + i h v) / h (h = 1e-30) is
exact up to round-off; here it is checked against a central difference
(R(U + e v) - R(U - e v)) / 2e, whose truncation error is O(e^2), on the
smooth branches of every flux and boundary set, with the limiter frozen as in
the reference's Jv (Solver.cpp:1366).
"""
import numpy as np
import pytest

from mycode_amd import cfd, configs
from oracle import oracle


def _setup(config, scale, flux):
    mesh, p, U = configs.setup(config, scale)
    p.flux = flux
    o = oracle.OracleSolver(oracle.OracleMesh.from_mesh(mesh), p)
    if config == 5:
        o.set_face_flux(configs.shock_band(mesh))
    rng = np.random.default_rng(config * 10 + flux)
    U = U.copy()
    U[:, :4] *= 1 + 0.01 * rng.standard_normal((U.shape[0], 4))
    # a small vertical velocity keeps |u_n| away from the kink at zero
    U[:, 2] += 0.01 * U[:, 1] * rng.standard_normal(U.shape[0])
    if p.nvar == 6:  # SST: perturbed freestream k and omega
        U[:, 4] *= 1 + 0.5 * rng.random(U.shape[0])
        U[:, 5] *= 1 + 0.5 * rng.random(U.shape[0])
    v = rng.standard_normal(U.shape) * np.abs(U).mean(0)
    v[:, p.nvar:] = 0
    return o, U, v, p


@pytest.mark.parametrize("config,scale", [(1, 0.1), (2, 0.05), (3, 0.03), (4, 0.03), (5, 0.02)])
@pytest.mark.parametrize("flux", [0, 1, 2])
def test_complex_step_matches_central_difference(config, scale, flux):
    o, U, v, p = _setup(config, scale, flux)
    o.compute_residuals(U, True)  # forms the limiter, then freeze it
    R, dR = o.residual_complex_step(U, v, 1e-30, update_limiters=False)
    assert np.array_equal(R, o.compute_residuals(U, False))  # real part is the residual
    e = 1e-7 * np.linalg.norm(U) / np.linalg.norm(v)
    Rp = o.compute_residuals(U + e * v, False)
    Rm = o.compute_residuals(U - e * v, False)
    cd = (Rp - Rm) / (2 * e)
    rel = np.linalg.norm(cd - dR) / np.linalg.norm(dR)
    assert rel < 1e-5, rel


def test_complex_step_is_linear_in_direction():
    o, U, v, p = _setup(2, 0.05, 1)
    o.compute_residuals(U, True)
    _, d1 = o.residual_complex_step(U, v, 1e-30, False)
    _, d2 = o.residual_complex_step(U, 3.0 * v, 1e-30, False)
    assert np.linalg.norm(d2 - 3 * d1) <= 1e-12 * np.linalg.norm(d2)
