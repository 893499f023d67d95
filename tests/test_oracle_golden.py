"""The oracle (CPU restatement) against the reference's own outputs.

tests/golden holds vectors written by the compiled reference
(tests/golden/gen_golden.py).  Every check here is bit-exact: the oracle keeps
the reference's operation order and is compiled with its flags.  When
oracle/_ref/libref.so exists (this build container), the live reference is
also compared on the shipped meshes.
"""
import os

import numpy as np
import pytest

from helpers import bump_params
from mycode_amd import cfd
from oracle import oracle

MESHES = ["bump_1230", "pm_1x", "laminar_flat_plate", "supersonic_wedge_flow", "turbulent_flat_plate"]


def _omesh(g):
    return oracle.OracleMesh(g["x"], g["y"], g["cell_ptr"], g["cell_nodes"], g["bedges"])


def test_mesh_matches_reference(bump_golden):
    g = bump_golden
    m = _omesh(g)
    f, c = m.faces(), m.cells()
    for k, gk in (("owner", "f_owner"), ("neighbor", "f_neighbor"), ("bc_tag", "f_tag"), ("node1", "f_node1"),
                  ("node2", "f_node2"), ("nx", "f_nx"), ("ny", "f_ny"), ("area", "f_area"), ("x_mid", "f_xm"),
                  ("y_mid", "f_ym")):
        assert np.array_equal(f[k], g[gk]), k
    for k, gk in (("volume", "c_vol"), ("x_c", "c_xc"), ("y_c", "c_yc")):
        assert np.array_equal(c[k], g[gk]), k


@pytest.mark.parametrize("idx", range(12))
def test_residual_all_schemes_bitexact(bump_golden, idx):
    g = bump_golden
    fl, sp = g["schemes"][idx]
    s = oracle.OracleSolver(_omesh(g), bump_params(g, fl, sp))
    R = s.compute_residuals(g["U"])
    assert np.array_equal(R[:, :4], g["R_schemes"][idx])
    assert not R[:, 4:].any()


def test_gradients_limiter_dt_precond_bitexact(bump_golden):
    g = bump_golden
    s = oracle.OracleSolver(_omesh(g), bump_params(g, 2, 2))
    s.compute_residuals(g["U"])
    assert np.array_equal(s.gradients(), g["grad_hllc_lim"])
    assert np.array_equal(s.limiters(), g["phi_hllc_lim"])
    s.compute_local_time_step(g["U"], 5.0)
    assert np.array_equal(s.dt_local(), g["dt_cfl5"])
    assert np.array_equal(s.diagonal_preconditioner(g["U"]), g["D_cfl5"])


@pytest.mark.parametrize("fs", ["00", "11", "22"])
def test_fd_jv_and_gmres_bitexact(bump_golden, fs):
    g = bump_golden
    s = oracle.OracleSolver(_omesh(g), bump_params(g, int(fs[0]), int(fs[1])))
    R = s.compute_residuals(g["U"])
    s.compute_local_time_step(g["U"], 5.0)
    assert np.array_equal(s.apply_jacobian(g["U"], R, g["v"])[:, :4], g[f"Jv_{fs}"])
    x, it = s.solve_gmres(g["U"], R, -R)
    assert it == int(g[f"gmres_it_{fs}"])
    assert np.array_equal(x[:, :4], g[f"gmres_x_{fs}"])


@pytest.mark.parametrize("fs", ["00", "22"])
def test_newton_trajectory_bitexact(bump_golden, fs):
    g = bump_golden
    s = oracle.OracleSolver(_omesh(g), bump_params(g, int(fs[0]), int(fs[1])))
    U, info = s.solve_implicit(g["U0"], tol=1e-12, max_steps=int(g[f"newton_{fs}_steps"]))
    assert np.array_equal(U[:, :4], g[f"newton_{fs}_U"])


def test_explicit_step_bitexact(bump_golden):
    g = bump_golden
    s = oracle.OracleSolver(_omesh(g), bump_params(g, 2, 2))
    assert np.array_equal(s.time_step_explicit(g["U"], 1e-6)[:, :4], g["explicit_U"])


def test_synthetic_wedge_via_reference_reader(wedge_golden):
    """Generator-written Gmsh read back by the reference: same mesh, same residual."""
    g = wedge_golden
    m = _omesh(g)
    assert np.array_equal(m.faces()["nx"], g["f_nx"])
    for fl, sp in ((2, 2), (0, 0)):
        p = cfd.default_params()
        cfd.flow_conditions(p, 2.5, 101325.0, 300.0)
        p.flux, p.spatial, p.bcset, p.nvar = fl, sp, cfd.BC_BUMP, 4
        s = oracle.OracleSolver(m, p)
        assert np.array_equal(s.compute_residuals(g["U"])[:, :4], g[f"R_{fl}{sp}"])


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (needs /root/reference)")
@pytest.mark.parametrize("name", MESHES)
def test_live_reference_on_shipped_meshes(name, tmp_path):
    src = os.path.join(oracle.REFERENCE, "mesh", name)
    lf = tmp_path / name
    lf.write_bytes(open(src, "rb").read().replace(b"\r", b""))
    rm = oracle.RefMesh(str(lf))
    om = oracle.OracleMesh(gmsh=src)  # the oracle strips CR itself
    rf, of = rm.faces(), om.faces()
    for k in rf:
        assert np.array_equal(rf[k], of[k]), k
    rng = np.random.default_rng(7)
    for fl, sp in ((1, 1), (2, 2)):
        rs = oracle.RefSolver(rm, 0.5, 101325.0, 300.0, fl, sp)
        p = cfd.default_params()
        cfd.flow_conditions(p, 0.5, 101325.0, 300.0)
        p.flux, p.spatial, p.bcset, p.nvar = fl, sp, cfd.BC_BUMP, 4
        os_ = oracle.OracleSolver(om, p)
        U = rs.initialize()
        U[:, :4] *= 1 + 0.01 * rng.standard_normal((U.shape[0], 4))
        assert np.array_equal(rs.compute_residuals(U), os_.compute_residuals(U))
        assert np.array_equal(rs.gradients(), os_.gradients())
