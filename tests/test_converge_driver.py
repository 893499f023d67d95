"""Control flow of adjoint.converge (host logic, no GPU): the pseudo-time
phase stops at the Newton threshold, the Newton steps take the solve to the
target, and a failed Newton step hands the state back to the pseudo-time
iteration with the original target as its absolute tolerance.  The solver
and the Newton stage are stand-ins that record their calls."""
import numpy as np
import pytest

from mycode_amd import adjoint, cfd


class FakeSolver:
    def __init__(self, vm0=1e8, pseudo_final=None):
        self.vm0 = vm0
        self.calls = []
        self.pseudo_final = pseudo_final  # vm each pseudo-time call ends at (list), else the target it was given

    def gmres_opts(self, **kw):
        return kw

    def newton_opts(self, gmres=None, **kw):
        return dict(kw, gmres=gmres)

    def solve_implicit(self, U, o):
        self.calls.append(o)
        start = self.vm0 if len(self.calls) == 1 else 1.0
        target = max(o["rel_drop"] * start, o["tolerance"])
        end = self.pseudo_final.pop(0) if self.pseudo_final else target
        r = cfd.NewtonResult()
        r.steps, r.converged, r.res_vm0, r.res_vm, r.seconds, r.total_gmres_iters = 10, int(end <= target), start, end, 1.0, 100
        hist = {"vm": np.array([start, end]), "iters": np.array([100, 0]), "cfl": np.array([1.0, 50.0])}
        return U + 1.0, r, hist


def test_newton_stage_finishes_the_solve(monkeypatch):
    seen = {}

    def fake_newton(solver, U, target, **kw):
        seen["target"] = target
        return U + 10.0, 3, 700, [1e-2, 1e-3, 0.5 * target], True

    monkeypatch.setattr(adjoint, "newton_fgmres", fake_newton)
    s = FakeSolver()
    U, res, hist = adjoint.converge(s, np.zeros((4, 6)), rel_drop=1e-12, newton_below=1e-4)
    assert len(s.calls) == 1 and s.calls[0]["rel_drop"] == 1e-4
    assert seen["target"] == pytest.approx(1e-12 * s.vm0)
    assert res.converged == 1 and res.steps == 13 and res.total_gmres_iters == 800
    assert res.res_vm == pytest.approx(0.5e-12 * s.vm0)
    assert list(hist["vm"][-3:]) == [1e-2, 1e-3, 0.5e-12 * s.vm0]
    assert np.all(U == 11.0)


def test_failed_newton_step_resumes_pseudo_time(monkeypatch):
    def fake_newton(solver, U, target, **kw):
        return U + 10.0, 2, 500, [5.0, 4.0], False

    monkeypatch.setattr(adjoint, "newton_fgmres", fake_newton)
    s = FakeSolver()
    U, res, hist = adjoint.converge(s, np.zeros((4, 6)), rel_drop=1e-12, newton_below=1e-4, max_steps=600)
    assert len(s.calls) == 2
    second = s.calls[1]
    assert second["rel_drop"] == 0.0 and second["tolerance"] == pytest.approx(1e-12 * s.vm0)
    assert second["max_steps"] == 600 - 12 and second["cfl0"] == 50.0
    assert res.steps == 22 and res.total_gmres_iters == 700 and res.converged == 1
    assert res.res_vm0 == s.vm0
    assert np.all(U == 12.0)


def test_straight_newton_and_off_switch(monkeypatch):
    calls = []

    def fake_newton(solver, U, target, **kw):
        calls.append(target)
        return U, 1, 10, [0.5 * target], True

    monkeypatch.setattr(adjoint, "newton_fgmres", fake_newton)
    s = FakeSolver(pseudo_final=[1e3])  # a re-solve: the pseudo-time phase stops at once
    adjoint.converge(s, np.zeros((4, 6)), rel_drop=0.0, tolerance=2.0, newton_below=1.0)
    assert s.calls[0]["rel_drop"] == 1.0 and calls == [2.0]
    s2 = FakeSolver()
    adjoint.converge(s2, np.zeros((4, 6)), rel_drop=1e-10, newton_below=0.0)
    assert len(s2.calls) == 1 and s2.calls[0]["rel_drop"] == 1e-10 and calls == [2.0]


def test_newton_steps_short_of_the_target_resume_pseudo_time(monkeypatch):
    def fake_newton(solver, U, target, **kw):
        return U + 10.0, 12, 900, [2.0 ** -k for k in range(12)], True  # halving, but the step cap is reached

    monkeypatch.setattr(adjoint, "newton_fgmres", fake_newton)
    s = FakeSolver()
    U, res, hist = adjoint.converge(s, np.zeros((4, 6)), rel_drop=1e-12, newton_below=1e-4, max_steps=600)
    assert len(s.calls) == 2 and s.calls[1]["tolerance"] == pytest.approx(1e-12 * s.vm0)
    assert res.converged == 1 and res.steps == 10 + 12 + 10
