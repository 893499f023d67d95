"""Line-implicit symmetric block Gauss-Seidel preconditioner.

Bars: one application equals the same sweeps done in numpy with dense
per-segment solves from the exported blocks (1e-9 relative in FP64: the
device eliminates block-tridiagonally, numpy solves densely, and the blocks
mix units so their condition numbers are large; 1e-4 in FP32); on the SST
plate at a high CFL, GMRES with it needs fewer iterations than with the
point block SGS and still meets its tolerance.  (With the FD product
instead of the tangent, GMRES stalls at its cap there with either
preconditioner: the FD step is set by |U|, which the energy dominates, so
the k and omega directions are differenced poorly.)
"""
import numpy as np
import pytest

from mycode_amd import cfd, configs

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


def _line_sweeps(B, A, nbr, sp, cells, cp, x):
    N, nv = x.shape
    z = np.zeros_like(x)
    nc = len(cp) - 1
    for col in list(range(nc)) + list(range(nc - 2, -1, -1)):
        for k in range(cp[col], cp[col + 1]):
            seg = cells[sp[k]:sp[k + 1]]
            L = len(seg)
            pos = {int(c): i for i, c in enumerate(seg)}
            T = np.zeros((L * nv, L * nv))
            r = np.zeros(L * nv)
            for i, c in enumerate(seg):
                T[i * nv:(i + 1) * nv, i * nv:(i + 1) * nv] = B[c]
                r[i * nv:(i + 1) * nv] = x[c]
                for q in range(4):
                    o = int(nbr[c, q])
                    if o < 0:
                        continue
                    j = pos.get(o)
                    if j is not None:
                        assert abs(j - i) == 1, "segment with a non-consecutive face neighbour"
                        T[i * nv:(i + 1) * nv, j * nv:(j + 1) * nv] = A[c, q]
                    else:
                        r[i * nv:(i + 1) * nv] -= A[c, q] @ z[o]
            z[seg] = np.linalg.solve(T, r).reshape(L, nv)
    return z


@pytest.mark.parametrize("config,scale", [(3, 0.02), (2, 0.05)])
@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("split", [False, True])
def test_line_apply_matches_dense_sweeps(config, scale, prec, split, monkeypatch):
    # small meshes sweep each color in one launch; large ones run the
    # single-cell segments in a kernel of their own (forced here by the
    # CFD_LINE_FUSED_MAX_CELLS threshold, read when the context is created)
    monkeypatch.setenv("CFD_LINE_FUSED_MAX_CELLS", "0" if split else "262144")
    mesh, p, U = _state(config, scale)
    s = cfd.Solver(mesh, p)
    nv = p.nvar
    s.compute_residuals(U, True)
    s.compute_local_time_step(U, 1e3)
    rng = np.random.default_rng(17)
    x = rng.standard_normal(U.shape)
    x[:, nv:] = 0
    z = s.line_precond(U, x, prec)
    _, A, nbr, _ = s.block_sgs(U, blocks=True)
    B, _ = s.block_jacobi(U, 64)
    sp, cells, cp = mesh.lines()
    zr = _line_sweeps(B, A, nbr, sp, cells, cp, x[:, :nv])
    err = np.linalg.norm(z[:, :nv] - zr) / np.linalg.norm(zr)
    print(f"config={config} FP{prec} split={split} N={s.N} segments={len(sp) - 1} longest={np.diff(sp).max()} "
          f"colors={len(cp) - 1} error {err:.2e}", flush=True)
    assert err < (1e-9 if prec == 64 else 1e-4)


@pytest.mark.parametrize("jv,prec", [(1, 64), (1, 32)])
def test_gmres_line_on_the_sst_plate(jv, prec):
    mesh, p, U = _state(3, 0.1, seed=3)
    s = cfd.Solver(mesh, p)
    R = s.compute_residuals(U)
    s.compute_local_time_step(U, 1e3)
    its = {}
    for pc in (2, 3):
        x, its[pc] = s.solve_gmres(U, R, -R, s.gmres_opts(m=60, tol=1e-2, jv=jv, precision=prec, precond=pc))
    print(f"jv={jv} FP{prec} CFL 1e3: GMRES iterations block SGS {its[2]}, line SGS {its[3]}", flush=True)
    assert its[3] < its[2] and its[3] < 60
    Jx = s.jv_tangent(U, x, 64)
    assert np.linalg.norm(Jx + R) <= 2e-2 * np.linalg.norm(R)
