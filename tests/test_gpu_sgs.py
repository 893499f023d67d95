"""Multicolor block symmetric Gauss-Seidel preconditioner.

Bars: coupling blocks of the first-order inviscid Jacobian against the
oracle's complex-step derivative (1e-12 relative); the assembled first-order
Jacobian (diagonal + coupling blocks) times a vector against the GPU tangent
(1e-12 relative); the GPU sweeps against the same sweeps in numpy from the
exported blocks (1e-12 FP64, 1e-5 FP32); M z = x for the SGS factors
(1e-9, round-off amplified by the block condition numbers); the transposed
solve by the dot-product identity (1e-12).
"""
import numpy as np
import pytest

from mycode_amd import cfd, configs
from oracle import oracle

pytestmark = pytest.mark.gpu


def _state(config, scale, seed=5):
    mesh, p, U = configs.setup(config, scale)
    rng = np.random.default_rng(seed)
    U = U.copy()
    U[:, :4] *= 1 + 0.01 * rng.standard_normal((U.shape[0], 4))
    U[:, 2] += 0.01 * U[:, 1] * rng.standard_normal(U.shape[0])
    if p.nvar == 6:
        U[:, 4] *= 1 + 0.5 * rng.random(U.shape[0])
        U[:, 5] *= 1 + 0.5 * rng.random(U.shape[0])
    else:
        U[:, 4:] = 0
    return mesh, p, U


def _first_order(mesh, p):
    p1 = p.copy()
    p1.spatial, p1.viscous = cfd.FIRST_ORDER, 0
    if p1.nvar == 6:
        p1.turb = 0
        p1.nvar = 4
    return p1


def _matvec(B, A, nbr, v):
    y = np.einsum("nij,nj->ni", B, v)
    for q in range(4):
        idx = np.where(nbr[:, q] >= 0, nbr[:, q], 0)
        y += np.einsum("nij,nj->ni", A[:, q], v[idx])
    return y


def _sweeps(Bi, A, nbr, col, x, transpose=False):
    """z = M^-1 x (or M^-T x) by the same color sweeps, in numpy."""
    N, nv = x.shape
    z = np.zeros_like(x)
    nc = int(col.max()) + 1
    order = list(range(nc)) + list(range(nc - 2, -1, -1))
    if transpose:  # back slot of each coupling
        back = np.full((N, 4), -1)
        for c in range(N):
            for q in range(4):
                o = nbr[c, q]
                if o >= 0:
                    back[c, q] = int(np.where(nbr[o] == c)[0][0])
    for cc in order:
        cells = np.where(col == cc)[0]
        s = x[cells].copy()
        for q in range(4):
            o = nbr[cells, q]
            ok = o >= 0
            oc = np.where(ok, o, 0)
            if transpose:
                r = np.where(ok, back[cells, q], 0)
                blk = np.transpose(A[oc, r], (0, 2, 1))
            else:
                blk = A[cells, q]
            s -= np.where(ok[:, None], np.einsum("nij,nj->ni", blk, z[oc]), 0.0)
        Bc = np.transpose(Bi[cells], (0, 2, 1)) if transpose else Bi[cells]
        z[cells] = np.einsum("nij,nj->ni", Bc, s)
    return z


@pytest.mark.parametrize("config,scale", [(1, 0.1), (2, 0.05), (4, 0.03)])
def test_coupling_blocks(config, scale):
    mesh, p, U = _state(config, scale)
    p1 = _first_order(mesh, p)
    nv = p1.nvar
    s1 = cfd.Solver(mesh, p1)
    o1 = oracle.OracleSolver(oracle.OracleMesh.from_mesh(mesh), p1)
    s1.compute_residuals(U, True)
    s1.compute_local_time_step(U, 3.0)
    _, A, nbr, col = s1.block_sgs(U, blocks=True)
    print(f"config={config} N={s1.N} colors={col.max() + 1}", flush=True)
    # a distance-1 coloring and a symmetric neighbor table
    for q in range(4):
        ok = nbr[:, q] >= 0
        assert np.all(col[ok] != col[nbr[ok, q]])
        for c in np.where(ok)[0][:200]:
            assert c in nbr[nbr[c, q]]
    rng = np.random.default_rng(config)
    for c in rng.choice(s1.N, 5, replace=False):
        for q in range(4):
            o = nbr[c, q]
            if o < 0:
                continue
            for k in range(nv):
                e = np.zeros_like(U)
                e[o, k] = 1.0
                _, dR = o1.residual_complex_step(U, e, 1e-30, True)
                assert np.allclose(A[c, q, :, k], dR[c, :nv], rtol=1e-12,
                                   atol=1e-12 * max(np.abs(dR[c, :nv]).max(), 1e-300))
    # the whole first-order Jacobian: diagonal + coupling blocks against the tangent
    B, _ = s1.block_jacobi(U, 64)
    v = rng.standard_normal(U.shape) * np.abs(U).mean(0)
    v[:, nv:] = 0
    Jv = s1.jv_tangent(U, v, 64)[:, :nv]
    y = _matvec(B, A, nbr, v[:, :nv])
    assert np.linalg.norm(y - Jv) <= 1e-12 * np.linalg.norm(Jv)


@pytest.mark.parametrize("config,scale", [(2, 0.05), (3, 0.02), (4, 0.03)])
@pytest.mark.parametrize("prec", [64, 32])
def test_sgs_apply(config, scale, prec):
    mesh, p, U = _state(config, scale)
    s = cfd.Solver(mesh, p)
    nv = p.nvar
    s.compute_residuals(U, True)
    s.compute_local_time_step(U, 50.0)
    rng = np.random.default_rng(11)
    x = rng.standard_normal(U.shape)
    x[:, nv:] = 0
    z, A, nbr, col = s.block_sgs(U, x, precision=prec, blocks=True)
    B, Bi = s.block_jacobi(U, prec)
    zr = _sweeps(Bi, A, nbr, col, x[:, :nv])
    err = np.linalg.norm(z[:, :nv] - zr) / np.linalg.norm(zr)
    print(f"config={config} FP{prec} N={s.N} colors={col.max() + 1} sweep error {err:.2e}", flush=True)
    assert err < (1e-12 if prec == 64 else 1e-5)
    if prec == 64:
        # M z = x with M = (D + L) D^-1 (D + U)
        lower = np.zeros_like(A)
        upper = np.zeros_like(A)
        for q in range(4):
            o = nbr[:, q]
            ok = o >= 0
            hi = ok & (col[np.where(ok, o, 0)] > col)
            lo = ok & ~hi
            upper[hi, q] = A[hi, q]
            lower[lo, q] = A[lo, q]
        zz = z[:, :nv]
        y = _matvec(B, upper, nbr, zz)
        w = np.einsum("nij,nj->ni", Bi, y)
        r = _matvec(B, lower, nbr, w)
        assert np.linalg.norm(r - x[:, :nv]) <= 1e-9 * np.linalg.norm(x[:, :nv])
        # transposed solve: <M^-1 a, b> = <a, M^-T b>
        b = rng.standard_normal(U.shape)
        b[:, nv:] = 0
        zt, _, _, _ = s.block_sgs(U, b, precision=64, transpose=True)
        lhs, rhs = float(np.sum(z * b)), float(np.sum(x * zt))
        assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), np.linalg.norm(z) * np.linalg.norm(b))
        ztr = _sweeps(Bi, A, nbr, col, b[:, :nv], transpose=True)
        assert np.linalg.norm(zt[:, :nv] - ztr) <= 1e-12 * np.linalg.norm(ztr)


@pytest.mark.parametrize("jv,prec", [(0, 64), (1, 64), (1, 32), (2, 64)])
def test_gmres_block_sgs(jv, prec):
    mesh, p, U = _state(2, 0.1, seed=3)
    s = cfd.Solver(mesh, p)
    R = s.compute_residuals(U)
    # the transposed operator carries information against the flow; at the
    # same CFL it needs more iterations, so it is tested at a smaller one
    cfl = 20.0 if jv == 2 else 200.0
    s.compute_local_time_step(U, cfl)
    its = {}
    for pc in (1, 2):
        x, its[pc] = s.solve_gmres(U, R, -R, s.gmres_opts(m=60, tol=1e-3, jv=jv, precision=prec, precond=pc))
    print(f"jv={jv} FP{prec} CFL {cfl}: GMRES iterations block-Jacobi {its[1]}, block SGS {its[2]}", flush=True)
    assert its[2] <= its[1] and its[2] < 60
    Ax = s.vjp_adjoint(U, x) if jv == 2 else s.jv_tangent(U, x, 64)
    assert np.linalg.norm(Ax + R) <= 3e-3 * np.linalg.norm(R)
