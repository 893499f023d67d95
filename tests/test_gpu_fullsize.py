"""Parity at the configurations' full sizes (the other GPU tests run the same
paths on coarsened meshes).

  * configuration 3 (SST plate, 1,320,146 cells), 4 (bump, 824,378 cells)
    and 5 (SWBLI, 4,399,865 cells): the HIP residual against the oracle on
    a seeded perturbed state.  Bars as in test_gpu_residual.py: bit-exact on
    the SWBLI set (no `pow`, no SST transcendental functions); within 1e-13
    of the cell flux scale S_i = sum |F_f A_f| where `pow` (bump inlet) or
    the SST tanh / sqrt may differ by an ulp between ROCm and glibc;
  * configurations 3 and 5: the FP64 tangent against the oracle's complex
    step, 1e-11 relative (both exact derivatives on the active branch);
  * configuration 3: the adjoint's dot-product identity
    <w, J v> = <J^T w, v> to 1e-12 of |w| |J v| (as on the coarse meshes),
    and run-to-run bitwise reproducibility of the residual.
"""
import numpy as np
import pytest

from mycode_amd import cfd, configs
from oracle import oracle

pytestmark = pytest.mark.gpu


def norm_err(R, Rref, S):
    S = np.where(S > 0, S, 1.0)
    return float(np.max(np.abs(R - Rref)[:, : S.shape[1]] / S))


def _state(config, mesh, p, U0):
    rng = np.random.default_rng(100 + config)
    U = U0.copy()
    U[:, :4] *= 1 + 0.01 * rng.standard_normal((U.shape[0], 4))
    U[:, 2] += 0.01 * U[:, 1] * rng.standard_normal(U.shape[0])
    if p.nvar == 6:
        U[:, 4] *= 1 + 0.5 * rng.random(U.shape[0])
        U[:, 5] *= 1 + 0.5 * rng.random(U.shape[0])
    else:
        U[:, 4:] = 0.0
    return U


_CACHE = {}


def _case(config):
    if config not in _CACHE:
        _CACHE.clear()  # one full-size case resident at a time
        mesh, p, U0 = configs.setup(config, 1.0)
        U = _state(config, mesh, p, U0)
        o = oracle.OracleSolver(oracle.OracleMesh.from_mesh(mesh), p)
        s = cfd.Solver(mesh, p, assembly=1)
        if config == 5:
            alt = configs.shock_band(mesh)
            o.set_face_flux(alt)
            s.set_face_flux(alt)
        _CACHE[config] = (mesh, p, U, o, s)
    return _CACHE[config]


@pytest.mark.parametrize("config", [3, 4, 5])
def test_fullsize_residual_vs_oracle(config):
    mesh, p, U, o, s = _case(config)
    print(f"full-size residual: configuration {config}, {mesh.n_cells} cells", flush=True)
    R = s.compute_residuals(U)
    Ro = o.compute_residuals(U)
    if p.bcset != cfd.BC_BUMP and not p.turb:
        assert np.array_equal(R, Ro)
    else:
        S = o.residual_scale(U)[:, : p.nvar]
        assert norm_err(R, Ro, S) <= 1e-13
    assert np.array_equal(R, s.compute_residuals(U))


@pytest.mark.parametrize("config", [3, 5])
def test_fullsize_tangent_vs_complex_step(config):
    mesh, p, U, o, s = _case(config)
    print(f"full-size tangent: configuration {config}, {mesh.n_cells} cells", flush=True)
    s.compute_residuals(U, True)
    o.compute_residuals(U, True)
    s.compute_local_time_step(U, 3.0)
    rng = np.random.default_rng(7)
    v = rng.standard_normal(U.shape) * np.abs(U).mean(0)
    v[:, p.nvar:] = 0
    J = s.jv_tangent(U, v, 64)
    _, dR = o.residual_complex_step(U, v, 1e-30, update_limiters=False)
    Jref = dR + v * (mesh.cells()["volume"] / (s.dt_local() + 1e-30))[:, None]
    rel = float(np.linalg.norm(J - Jref) / np.linalg.norm(Jref))
    assert rel < 1e-11, rel


def test_fullsize_adjoint_dot_product_identity():
    mesh, p, U, o, s = _case(3)
    print(f"full-size adjoint identity: configuration 3, {mesh.n_cells} cells", flush=True)
    s.compute_residuals(U, True)
    s.compute_local_time_step(U, 3.0)
    rng = np.random.default_rng(9)
    v = rng.standard_normal(U.shape) * np.abs(U).mean(0)
    w = rng.standard_normal(U.shape)
    v[:, p.nvar:] = 0
    w[:, p.nvar:] = 0
    Jv = s.jv_tangent(U, v, 64)
    JTw = s.vjp_adjoint(U, w)
    a, b = float(np.sum(w * Jv)), float(np.sum(JTw * v))
    assert abs(a - b) <= 1e-12 * np.linalg.norm(w) * np.linalg.norm(Jv), (a, b)
