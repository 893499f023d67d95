"""The HIP path against the reference's commented-in variants (fixtures in
tests/golden/ref_variants.npz, written by the reference compiled with the
edits of oracle/ref_variant.sh; see test_oracle_variants.py).

Parity bars, each written below:
  * residual, buffer assembly: bit-exact on the flat-plate and supersonic
    sets (no `pow` on those paths); the viscous bump set uses the isentropic
    inlet (pow, Solver.cpp:820), where the ROCm and glibc pow may differ by
    an ulp, so its bar is 1e-13 of the cell's flux scale S_i;
  * local time step bit-exact, diagonal preconditioner within 1e-14, FD Jv
    within 1e-7 (the device norms that set eps sum in another order);
  * explicit steps within 1e-14 relative per variable;
  * Newton trajectory of the reference controller, 10 steps: the state
    within 1e-8 of the reference's;
  * Cf and the wall-pressure rows against the reference's CSV rows at their
    6 significant digits; calculateDrag bit-exact.
"""
import numpy as np
import pytest

from conftest import load_golden
from mycode_amd import cfd, configs
from oracle import oracle
from test_oracle_variants import PLATE, SUPER, VISC, _U, direction, variant_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gv():
    return load_golden("ref_variants.npz")


def _mesh(g, pre):
    return cfd.Mesh.from_arrays(g[f"{pre}_x"], g[f"{pre}_y"], g[f"{pre}_cell_ptr"], g[f"{pre}_cell_nodes"],
                                g[f"{pre}_bedges"])


def _nu(g, pre):
    return float(g["plate_nu"]) if pre == "plate" else 0.01


def _gpu(g, pre, kind, fs):
    p = variant_params(kind["bcset"], kind["viscous"], kind["M"], _nu(g, pre), int(fs[0]), int(fs[1]))
    return cfd.Solver(_mesh(g, pre), p), p


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("pre,kind,fs", [("plate", PLATE, "11"), ("plate", PLATE, "22"), ("super", SUPER, "22"),
                                         ("super", SUPER, "00"), ("visc", VISC, "11"), ("visc", VISC, "22")])
def test_variant_residual(gv, pre, kind, fs):
    print(f"variant residual {pre} {fs}", flush=True)
    s, p = _gpu(gv, pre, kind, fs)
    U = _U(gv, pre)
    R = s.compute_residuals(U)[:, :4]
    ref = gv[f"{pre}_R_{fs}"]
    if pre == "visc":
        o = oracle.OracleSolver(oracle.OracleMesh(gv["visc_x"], gv["visc_y"], gv["visc_cell_ptr"],
                                                  gv["visc_cell_nodes"], gv["visc_bedges"]), p)
        S = o.residual_scale(U)[:, :4]
        assert float(np.max(np.abs(R - ref) / np.where(S > 0, S, 1.0))) < 1e-13
    else:
        assert np.array_equal(R, ref)


def test_plate_dt_precond_jv(gv):
    s, p = _gpu(gv, "plate", PLATE, "11")
    U = _U(gv, "plate")
    R = s.compute_residuals(U)
    s.compute_local_time_step(U, 5.0)
    assert np.array_equal(s.dt_local(), gv["plate_dt_cfl5"])
    D = s.diagonal_preconditioner(U)
    assert np.max(np.abs(D / gv["plate_D_cfl5"] - 1)) < 1e-14
    J = s.apply_jacobian(U, R, direction(U))
    assert rel(J[:, :4], gv["plate_Jv_11"]) < 1e-7


def test_plate_newton_trajectory(gv):
    s, p = _gpu(gv, "plate", PLATE, "10")
    U0 = np.tile(configs.freestream_state(p), (gv["plate_U"].shape[0], 1))
    steps = int(gv["plate_newton_steps"])
    U, res, hist = s.solve_implicit(U0, s.newton_opts(tolerance=1e-12, max_steps=steps))
    assert res.steps == steps
    assert rel(U[:, :4], gv["plate_newton_U"]) < 1e-8


def test_supersonic_explicit_steps(gv):
    s, p = _gpu(gv, "super", SUPER, "22")
    U = _U(gv, "super")
    dt = float(gv["super_dt"])
    U1 = s.time_step_explicit(U, dt)
    sc = np.abs(U[:, :4]).max(0)
    assert np.max(np.abs(U1[:, :4] - gv["super_explicit_U1"]) / sc) < 1e-14
    U2 = s.time_step_explicit(0.75 * U + 0.25 * U1, dt)
    assert np.max(np.abs(U2[:, :4] - gv["super_explicit_U2"]) / sc) < 1e-14


def test_plate_skin_friction_vs_reference_csv(gv):
    s, p = _gpu(gv, "plate", PLATE, "11")
    _, Cf, _ = s.wall_outputs(_U(gv, "plate"))
    f = s.mesh.faces()
    w = np.where((f["neighbor"] < 0) & (f["bc_tag"] == 4))[0]
    order = np.argsort(f["x_mid"][w], kind="stable")
    rows = gv["plate_cf_rows"]
    assert len(rows) == len(w) > 50
    cf = Cf[w][order]
    assert np.all(np.abs(cf - rows[:, 1]) <= 5e-6 * np.abs(cf))


def test_drag_vs_reference(gv):
    s, p = _gpu(gv, "visc", dict(VISC, viscous=0), "22")
    assert s.calculate_drag(_U(gv, "visc")) == float(gv["bump_drag"])
    s, p = _gpu(gv, "super", SUPER, "22")
    assert s.calculate_drag(_U(gv, "super")) == float(gv["super_drag"])


@pytest.mark.parametrize("pre,kind", [("plate", PLATE), ("super", SUPER)])
def test_wall_pressure_rows_vs_reference_csv(gv, pre, kind):
    s, p = _gpu(gv, pre, kind, "11")
    rows = s.extract_wall_pressure(_U(gv, pre))
    ref = gv[f"{pre}_wp_rows"]
    assert rows.shape == ref.shape and len(ref) > 10
    assert np.all(np.abs(rows - ref) <= 5e-6 * np.abs(rows) + 1e-300)
