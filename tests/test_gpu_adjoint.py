"""Reverse-mode product J^T w on the GPU.

Pinned by the dot-product identity <w, J v> = <J^T w, v> against the forward
tangent, which is itself pinned against the oracle's complex-step derivative
(test_gpu_tangent.py).  Bar: 1e-12 relative to |w||Jv| (FP64; both sides are
exact derivatives, the difference is summation round-off).  A column check
against complex step (J^T e_c for a few cells) closes the loop.
"""
import numpy as np
import pytest

from mycode_amd import cfd, configs
from oracle import oracle

pytestmark = pytest.mark.gpu

CASES = [(1, 0.1), (2, 0.05), (3, 0.02), (4, 0.03), (5, 0.02)]


def _setup(config, scale, flux, spatial=None, tfo=None):
    mesh, p, U = configs.setup(config, scale)
    p.flux = flux
    if spatial is not None:
        p.spatial = spatial
    if tfo is not None:
        p.turb_first_order = tfo
    s = cfd.Solver(mesh, p)
    o = oracle.OracleSolver(oracle.OracleMesh.from_mesh(mesh), p)
    if config == 5:
        alt = configs.shock_band(mesh)
        s.set_face_flux(alt)
        o.set_face_flux(alt)
    rng = np.random.default_rng(config * 7 + flux)
    U = U.copy()
    U[:, :4] *= 1 + 0.01 * rng.standard_normal((U.shape[0], 4))
    U[:, 2] += 0.01 * U[:, 1] * rng.standard_normal(U.shape[0])
    if p.nvar == 6:  # SST: perturbed freestream k and omega
        U[:, 4] *= 1 + 0.5 * rng.random(U.shape[0])
        U[:, 5] *= 1 + 0.5 * rng.random(U.shape[0])
    else:
        U[:, 4:] = 0
    return s, o, U, p, rng


@pytest.mark.parametrize("config,scale,tfo", [c + (None,) for c in CASES] + [(3, 0.02, 0), (3, 0.02, 1)])
@pytest.mark.parametrize("flux", [0, 1, 2])
def test_dot_product_identity(config, scale, tfo, flux):
    """Structured interior-face products (Euler, laminar, SST with first-order
    k and omega convection) and the general one (SST with k and omega
    reconstructed: turb_first_order = 0) against the forward tangent."""
    s, o, U, p, rng = _setup(config, scale, flux, tfo=tfo)
    print(f"config={config} flux={flux} turb_first_order={p.turb_first_order} N={s.N}", flush=True)
    s.compute_residuals(U, True)
    s.compute_local_time_step(U, 3.0)
    v = rng.standard_normal(U.shape) * np.abs(U).mean(0)
    w = rng.standard_normal(U.shape)
    v[:, p.nvar:] = 0
    w[:, p.nvar:] = 0
    Jv = s.jv_tangent(U, v, 64)
    JTw = s.vjp_adjoint(U, w)
    lhs, rhs = float(np.sum(w * Jv)), float(np.sum(JTw * v))
    scale_ = np.linalg.norm(w) * np.linalg.norm(Jv)
    assert abs(lhs - rhs) <= 1e-12 * scale_, (lhs, rhs)
    assert not JTw[:, p.nvar:].any()


@pytest.mark.parametrize("spatial", [0, 1, 2])
def test_adjoint_columns_vs_complex_step(spatial):
    """(J^T e_c)_d = (J e_d)_c: rows of the transpose against complex-step columns."""
    s, o, U, p, rng = _setup(2, 0.04, 1, spatial)
    s.compute_residuals(U, True)
    o.compute_residuals(U, True)
    s.compute_local_time_step(U, 3.0)
    dt = s.dt_local()
    vol = s.mesh.cells()["volume"]
    nv = p.nvar
    for c in rng.choice(s.N, 3, replace=False):
        for k in range(nv):
            e = np.zeros_like(U)
            e[c, k] = 1.0
            row = s.vjp_adjoint(U, e)  # row (c, k) of J
            # column (d, j) of J by complex step, compared on the stencil of c
            for d in rng.choice(s.N, 2, replace=False).tolist() + [int(c)]:
                for j in range(nv):
                    ed = np.zeros_like(U)
                    ed[d, j] = 1.0
                    _, dR = o.residual_complex_step(U, ed, 1e-30, False)
                    ref = dR[c, k] + (vol[c] / (dt[c] + 1e-30) if (c == d and j == k) else 0.0)
                    assert abs(row[d, j] - ref) <= 1e-11 * max(1.0, abs(ref)), (c, k, d, j)


def test_adjoint_deterministic():
    s, o, U, p, rng = _setup(4, 0.03, 2)
    s.compute_residuals(U, True)
    s.compute_local_time_step(U, 3.0)
    w = rng.standard_normal(U.shape)
    w[:, p.nvar:] = 0
    assert np.array_equal(s.vjp_adjoint(U, w), s.vjp_adjoint(U, w))


@pytest.mark.parametrize("config,scale,precisions", [(4, 0.05, (64, 32)), (1, 0.1, (64,))])
def test_transposed_gmres_fp32_and_fp64(config, scale, precisions):
    """GMRES on the transposed system (J^T + V/dtau) x = b with the block SGS
    preconditioner, with an FP64 and with an FP32 basis and preconditioner
    (the reverse product runs in FP64 in both): each solution's true
    residual, formed with the FP64 reverse product, meets the tolerance, and
    the two solutions agree to 1e-4 (the FP32 inner solve's error).  The
    design adjoint (configuration 4) uses the FP32 form inside its FP64
    defect correction.

    Configuration 1 runs the FP64 basis only: its shifted system at CFL 20
    has a condition number near 6.5e12 (the density adjoint sits six decades
    above the energy one), and an FP32 basis loses the solution at any
    tolerance (true residual ~1e2 |b| with the least-squares estimate at
    1e-6; the forward system in FP32 behaves the same).  With SST the omega
    rows span eleven decades more than the mean flow's, also beyond an FP32
    basis."""
    s, o, U, p, rng = _setup(config, scale, 2)
    print(f"config={config} N={s.N}", flush=True)
    s.compute_residuals(U, True)
    s.compute_local_time_step(U, 20.0)
    b = rng.standard_normal(U.shape)
    b[:, p.nvar:] = 0
    zero = np.zeros_like(U)
    xs = {}
    for prec in precisions:
        opts = s.gmres_opts(m=60, tol=1e-6, jv=2, precision=prec, precond=2, restarts=8)
        x, it = s.solve_gmres(U, zero, b, opts)
        res = np.linalg.norm(s.vjp_adjoint(U, x) - b) / np.linalg.norm(b)
        print(f"precision {prec}: {it} iterations, true relative residual {res:.2e}", flush=True)
        assert res < 1e-5
        xs[prec] = x
    if 32 not in xs:
        return
    assert np.linalg.norm(xs[32] - xs[64]) <= 1e-4 * np.linalg.norm(xs[64])
