"""GPU forward tangent J v against the oracle's complex-step derivative, the
finite-difference Jv of the reference, and FP32 against FP64.

Bars (stated per test): FP64 tangent vs complex step 1e-11 relative (both are
exact derivatives on the active branch; the difference is round-off in
sums that cancel); tangent vs the reference FD Jv 1e-5 relative (FD
truncation + round-off at eps ~ 1e-7 |U| / |v|); FP32 tangent vs FP64 tangent
3e-4 relative (single-precision values and derivatives, error stated in the
output).
"""
import numpy as np
import pytest

from mycode_amd import cfd, configs
from oracle import oracle

pytestmark = pytest.mark.gpu

CASES = [(1, 0.1), (2, 0.05), (3, 0.02), (4, 0.03), (5, 0.02)]


def _setup(config, scale, flux, assembly=0, tfo=None):
    mesh, p, U = configs.setup(config, scale)
    p.flux = flux
    if tfo is not None:
        p.turb_first_order = tfo
    o = oracle.OracleSolver(oracle.OracleMesh.from_mesh(mesh), p)
    s = cfd.Solver(mesh, p, assembly=assembly)
    if config == 5:
        alt = configs.shock_band(mesh)
        o.set_face_flux(alt)
        s.set_face_flux(alt)
    rng = np.random.default_rng(config * 10 + flux)
    U = U.copy()
    U[:, :4] *= 1 + 0.01 * rng.standard_normal((U.shape[0], 4))
    # a small vertical velocity keeps |u_n| away from the kink at zero
    U[:, 2] += 0.01 * U[:, 1] * rng.standard_normal(U.shape[0])
    if p.nvar == 6:  # SST: perturbed freestream k and omega
        U[:, 4] *= 1 + 0.5 * rng.random(U.shape[0])
        U[:, 5] *= 1 + 0.5 * rng.random(U.shape[0])
    else:
        U[:, 4:] = 0
    v = rng.standard_normal(U.shape) * np.abs(U).mean(0)
    v[:, p.nvar:] = 0
    return s, o, U, v, p


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("config,scale", CASES)
@pytest.mark.parametrize("flux", [0, 1, 2])
@pytest.mark.parametrize("assembly", [0, 1])
def test_tangent_fp64_vs_complex_step(config, scale, flux, assembly):
    s, o, U, v, p = _setup(config, scale, flux, assembly)
    print(f"config={config} flux={flux} assembly={assembly} N={s.N}", flush=True)
    R = s.compute_residuals(U, True)
    o.compute_residuals(U, True)
    s.compute_local_time_step(U, 3.0)
    dt = s.dt_local()
    vol = s.mesh.cells()["volume"]
    J = s.jv_tangent(U, v, 64)
    _, dR = o.residual_complex_step(U, v, 1e-30, update_limiters=False)
    Jref = dR + v * (vol / (dt + 1e-30))[:, None]
    assert rel(J, Jref) < 1e-11


@pytest.mark.parametrize("flux", [0, 1, 2])
def test_tangent_sst_reconstructed_kw_vs_complex_step(flux):
    """The SST plate with k and omega reconstructed like the mean flow
    (turb_first_order = 0; configuration 3 itself convects them first-order)."""
    s, o, U, v, p = _setup(3, 0.02, flux, 1, tfo=0)
    print(f"config=3 turb_first_order=0 flux={flux} N={s.N}", flush=True)
    s.compute_residuals(U, True)
    o.compute_residuals(U, True)
    s.compute_local_time_step(U, 3.0)
    dt = s.dt_local()
    vol = s.mesh.cells()["volume"]
    J = s.jv_tangent(U, v, 64)
    _, dR = o.residual_complex_step(U, v, 1e-30, update_limiters=False)
    Jref = dR + v * (vol / (dt + 1e-30))[:, None]
    assert rel(J, Jref) < 1e-11


@pytest.mark.parametrize("config,scale", CASES)
def test_tangent_vs_fd_jv(config, scale):
    s, o, U, v, p = _setup(config, scale, 1)
    R = s.compute_residuals(U, True)
    s.compute_local_time_step(U, 3.0)
    J = s.jv_tangent(U, v, 64)
    Jfd = s.apply_jacobian(U, R, v)
    assert rel(Jfd, J) < 1e-5


@pytest.mark.parametrize("config,scale", CASES)
def test_tangent_fp32_vs_fp64(config, scale):
    s, o, U, v, p = _setup(config, scale, 1)
    s.compute_residuals(U, True)
    s.compute_local_time_step(U, 3.0)
    J64 = s.jv_tangent(U, v, 64)
    J32 = s.jv_tangent(U, v, 32)
    err = rel(J32, J64)
    print(f"FP32 tangent relative error vs FP64: {err:.3e}")
    assert err < 3e-4


def test_tangent_linear_and_deterministic():
    s, o, U, v, p = _setup(2, 0.05, 1)
    s.compute_residuals(U, True)
    s.compute_local_time_step(U, 3.0)
    J1 = s.jv_tangent(U, v, 64)
    J2 = s.jv_tangent(U, 2.0 * v, 64)
    assert rel(J2, 2 * J1) < 1e-15
    assert np.array_equal(J1, s.jv_tangent(U, v, 64))


@pytest.mark.parametrize("prec", [64, 32])
def test_gmres_with_tangent(prec):
    mesh, p, U = configs.setup(2, 0.1)
    U = U.copy()
    U[:, :4] *= 1 + 0.01 * np.random.default_rng(3).standard_normal((U.shape[0], 4))
    s = cfd.Solver(mesh, p)
    R = s.compute_residuals(U)
    s.compute_local_time_step(U, 2.0)
    xfd, itfd = s.solve_gmres(U, R, -R, s.gmres_opts(m=30, tol=1e-3))
    xt, itt = s.solve_gmres(U, R, -R, s.gmres_opts(m=30, tol=1e-3, jv=1, precision=prec))
    assert abs(itt - itfd) <= 2
    assert rel(xt, xfd) < (1e-4 if prec == 64 else 1e-3)
    # the Newton correction solves the linear system to the requested tolerance
    Jx = s.jv_tangent(U, xt, 64)
    assert np.linalg.norm(Jx + R) <= 1.5e-3 * np.linalg.norm(R)


@pytest.mark.parametrize("config,scale", [(1, 0.1), (2, 0.05), (4, 0.03), (5, 0.02)])
def test_block_jacobi_blocks(config, scale):
    """Block-Jacobi: B Binv = I (1e-12 FP64, 1e-5 FP32), and with the viscous
    diagonal off, B - (V/dtau) I equals the oracle's complex-step diagonal
    block of the first-order residual (1e-12 relative)."""
    s, o, U, v, p = _setup(config, scale, 2)
    nv = p.nvar
    s.compute_residuals(U, True)
    s.compute_local_time_step(U, 3.0)
    # backward-error bar: |B Binv - I| <= c * eps * cond(B); the blocks mix SI
    # units (density ~1, energy ~1e5), so cond(B) is large
    for prec, eps in ((64, 2.2e-16), (32, 1.2e-7)):
        B, Bi = s.block_jacobi(U, prec)
        err = np.max(np.abs(B @ Bi - np.eye(nv)), axis=(1, 2))
        kappa = np.linalg.cond(B)
        assert np.all(err <= 50 * eps * kappa), (prec, float(np.max(err / (eps * kappa))))
    p1 = p.copy()
    p1.spatial, p1.viscous = cfd.FIRST_ORDER, 0
    s1 = cfd.Solver(s.mesh, p1)
    o1 = oracle.OracleSolver(oracle.OracleMesh.from_mesh(s.mesh), p1)
    if config == 5:
        alt = configs.shock_band(s.mesh)
        s1.set_face_flux(alt)
        o1.set_face_flux(alt)
    s1.compute_local_time_step(U, 3.0)
    B, _ = s1.block_jacobi(U, 64)
    tc = s.mesh.cells()["volume"] / (s1.dt_local() + 1e-30)
    rng = np.random.default_rng(9)
    for c in rng.choice(s.N, 6, replace=False):
        for k in range(nv):
            e = np.zeros_like(U)
            e[c, k] = 1.0
            _, dR = o1.residual_complex_step(U, e, 1e-30, True)
            col = B[c, :, k] - tc[c] * (np.arange(nv) == k)
            assert np.allclose(col, dR[c, :nv], rtol=1e-12, atol=1e-12 * np.abs(dR[c, :nv]).max())


@pytest.mark.parametrize("prec", [64, 32])
def test_gmres_block_jacobi(prec):
    mesh, p, U = configs.setup(2, 0.1)
    U = U.copy()
    U[:, :4] *= 1 + 0.01 * np.random.default_rng(3).standard_normal((U.shape[0], 4))
    s = cfd.Solver(mesh, p)
    R = s.compute_residuals(U)
    s.compute_local_time_step(U, 2.0)
    xd, itd = s.solve_gmres(U, R, -R, s.gmres_opts(m=40, tol=1e-3, jv=1, precision=64, precond=0))
    xb, itb = s.solve_gmres(U, R, -R, s.gmres_opts(m=40, tol=1e-3, jv=1, precision=prec, precond=1))
    print(f"GMRES iterations: diagonal {itd}, block-Jacobi FP{prec} {itb}")
    assert itb <= itd and itb < 40
    Jx = s.jv_tangent(U, xb, 64)
    assert np.linalg.norm(Jx + R) <= 2e-3 * np.linalg.norm(R)


@pytest.mark.parametrize("jv,prec", [(0, 64), (1, 64), (1, 32)])
def test_gmres_restarts_reach_tight_tolerance(jv, prec):
    mesh, p, U = configs.setup(2, 0.1)
    U = U.copy()
    U[:, :4] *= 1 + 0.01 * np.random.default_rng(4).standard_normal((U.shape[0], 4))
    s = cfd.Solver(mesh, p)
    R = s.compute_residuals(U)
    s.compute_local_time_step(U, 5.0)
    x, it = s.solve_gmres(U, R, -R, s.gmres_opts(m=20, tol=1e-8, jv=jv, precision=prec, precond=1, restarts=30))
    Jx = s.jv_tangent(U, x, 64)
    rel = np.linalg.norm(Jx + R) / np.linalg.norm(R)
    print(f"restarted GMRES jv={jv} FP{prec}: {it} iterations, relative residual {rel:.2e}")
    assert rel < (1e-6 if jv == 0 else 1e-7 if prec == 64 else 1e-5)
