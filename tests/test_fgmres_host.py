"""The flexible GMRES of mycode_amd.adjoint (the outer iteration of the
adjoint solve and of the Newton steps) on the CPU, with torch CPU tensors and
dense operators: its solutions against numpy's dense solve, with a fixed, a
varying (flexible) and no preconditioner, across restarts, from a nonzero
start, and the residual it reports against the one recomputed here."""
import numpy as np
import pytest
import torch

from mycode_amd.adjoint import _fgmres


def _system(n=120, seed=0):
    rng = np.random.default_rng(seed)
    A = np.eye(n) * 4.0 + rng.standard_normal((n, n)) / np.sqrt(n)
    A[np.arange(n - 1), np.arange(1, n)] += 1.5  # nonsymmetric
    b = rng.standard_normal(n)
    return A, b


def _run(A, b, m, precond, x0=None, tol=1e-12, max_outer=2000):
    n = len(b)
    At = torch.from_numpy(A)
    calls = []

    def apply_A(x, out):
        out.copy_(At @ x)

    def apply_M(v, z):
        calls.append(1)
        z.copy_(precond(v, len(calls)))
        return 3

    V = torch.zeros((m + 1, n), dtype=torch.float64)
    Z = torch.zeros((m, n), dtype=torch.float64)
    w = torch.zeros(n, dtype=torch.float64)
    x = torch.zeros(n, dtype=torch.float64) if x0 is None else torch.from_numpy(x0.copy())
    inner, outer, rel = _fgmres(apply_A, apply_M, torch.from_numpy(b), x, tol, V, Z, w, max_outer)
    return x.numpy(), inner, outer, rel, len(calls)


@pytest.mark.parametrize("m", [200, 15])
@pytest.mark.parametrize("kind", ["none", "jacobi", "varying"])
def test_fgmres_solves_dense_systems(m, kind):
    A, b = _system()
    d = np.diag(A).copy()
    rng = np.random.default_rng(5)

    def precond(v, k):
        if kind == "none":
            return v
        if kind == "jacobi":
            return v / torch.from_numpy(d)
        # a different preconditioner at every outer step: flexible GMRES must
        # still converge (plain right-preconditioned GMRES would not be valid)
        s = 1.0 + 0.5 * rng.standard_normal(len(d)) ** 2
        return v / torch.from_numpy(d * s)

    x, inner, outer, rel, calls = _run(A, b, m, precond)
    x_ref = np.linalg.solve(A, b)
    res = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
    assert rel <= 1e-12 and abs(res - rel) <= 1e-13
    assert np.linalg.norm(x - x_ref) <= 1e-10 * np.linalg.norm(x_ref)
    assert calls == outer and inner == 3 * outer


def test_fgmres_from_a_nonzero_start_and_stopping_at_the_tolerance():
    A, b = _system(seed=2)
    x_ref = np.linalg.solve(A, b)
    x0 = x_ref + 1e-3 * np.random.default_rng(9).standard_normal(len(b))
    x, _, outer_warm, rel, _ = _run(A, b, 50, lambda v, k: v, x0=x0, tol=1e-8)
    assert rel <= 1e-8
    _, _, outer_cold, _, _ = _run(A, b, 50, lambda v, k: v, tol=1e-8)
    assert outer_warm < outer_cold
    # a loose tolerance stops early, with the reported residual the true one
    x, _, outer, rel, _ = _run(A, b, 50, lambda v, k: v, tol=1e-2)
    assert 1e-4 < rel <= 1e-2
    assert abs(np.linalg.norm(b - A @ x) / np.linalg.norm(b) - rel) <= 1e-14


def test_fgmres_zero_rhs_and_no_product_of_a_zero_start():
    A, b = _system(seed=3)
    x, inner, outer, rel, calls = _run(A, np.zeros_like(b), 20, lambda v, k: v)
    assert (inner, outer, rel, calls) == (0, 0, 0.0, 0) and not np.any(x)
    seen = []
    At = torch.from_numpy(A)

    def apply_A(v, out):
        seen.append(float(torch.linalg.norm(v)))
        out.copy_(At @ v)

    n = len(b)
    V, Z, w = torch.zeros((21, n), dtype=torch.float64), torch.zeros((20, n), dtype=torch.float64), torch.zeros(n, dtype=torch.float64)
    _fgmres(apply_A, lambda v, z: (z.copy_(v), 0)[1], torch.from_numpy(b), torch.zeros(n, dtype=torch.float64),
            1e-6, V, Z, w, 100)
    assert seen and seen[0] > 0.0  # the zero initial guess is never multiplied
