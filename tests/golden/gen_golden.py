"""Write the golden fixtures of tests/golden from the compiled reference.

Runs in the build container only: it needs /root/reference and
oracle/_ref/libref.so (``make -C oracle ref``).  Every vector below is the
reference's own output (Solver.cpp compiled with its makefile flags) on:

  ref_bump_1230_th.npz   the reference mesh mesh/bump_1230_th (1,502 cells),
                         its active boundary conditions (Solver.cpp:806-863),
                         Mach 0.5 (main.cpp:63), all 4 x 3 flux / spatial
                         schemes, plus gradients, limiter, local time step,
                         diagonal preconditioner, FD Jv, GMRES and a Newton
                         trajectory;
  ref_wedge_q.npz        a seeded synthetic wedge mesh (configuration 1 at
                         quarter scale) written by cfd_mesh_write_gmsh and
                         read back by the reference's Mesh::load;
  ref_bump_1230_next.npz on the bump_1230_th mesh, Roe + second-order central:
                         two BDF1 dual-time steps (Solver::solveUnsteady, dt
                         1e-5) from a seeded perturbed state, the entropy
                         error it prints, the boundary-layer profile rows it
                         writes, and the SHA-256 of main.cpp's VTK output.

The shipped meshes have CRLF line ends, which the reference's readGmsh only
accepts in Windows text mode; a CR-stripped temporary copy is fed to it.
Usage:  python tests/golden/gen_golden.py
"""
from __future__ import annotations

import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from mycode_amd import cfd  # noqa: E402
from oracle import oracle  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
SCHEMES = [(fl, sp) for fl in (0, 1, 2, 3) for sp in (0, 1, 2)]


def lf_copy(src: str, tmp: str) -> str:
    dst = os.path.join(tmp, os.path.basename(src))
    with open(src, "rb") as f, open(dst, "wb") as g:
        g.write(f.read().replace(b"\r", b""))
    return dst


def mesh_arrays(path: str, rm: oracle.RefMesh) -> dict:
    pm = cfd.Mesh.read_gmsh(path)
    x, y, cp, cn, be = pm.desc()
    f, c = rm.faces(), rm.cells()
    return dict(x=x, y=y, cell_ptr=cp, cell_nodes=cn, bedges=be,
                f_owner=f["owner"], f_neighbor=f["neighbor"], f_tag=f["bc_tag"], f_node1=f["node1"],
                f_node2=f["node2"], f_nx=f["nx"], f_ny=f["ny"], f_area=f["area"], f_xm=f["x_mid"],
                f_ym=f["y_mid"], c_vol=c["volume"], c_xc=c["x_c"], c_yc=c["y_c"])


def gen_bump(tmp: str):
    src = os.path.join(oracle.REFERENCE, "mesh", "bump_1230_th")
    path = lf_copy(src, tmp)
    rm = oracle.RefMesh(path)
    out = mesh_arrays(path, rm)
    rng = np.random.default_rng(20240601)
    rs0 = oracle.RefSolver(rm, 0.5, 101325.0, 300.0, 2, 2)
    U0 = rs0.initialize()
    U = U0.copy()
    U[:, :4] *= 1.0 + 0.02 * rng.standard_normal((U.shape[0], 4))
    v = rng.standard_normal(U.shape)
    v[:, 4:] = 0.0
    out.update(M=0.5, P=101325.0, T=300.0, U0=U0, U=U, v=v)
    R_all = []
    for fl, sp in SCHEMES:
        rs = oracle.RefSolver(rm, 0.5, 101325.0, 300.0, fl, sp)
        R = rs.compute_residuals(U)
        R_all.append(R[:, :4])
        if (fl, sp) == (2, 2):
            out["grad_hllc_lim"] = rs.gradients()
            out["phi_hllc_lim"] = rs.limiters()
            rs.compute_local_time_step(U, 5.0)
            out["dt_cfl5"] = rs.dt_local()
            out["D_cfl5"] = rs.diagonal_preconditioner(U)
        if (fl, sp) in ((0, 0), (1, 1), (2, 2)):
            rs.compute_local_time_step(U, 5.0)
            out[f"Jv_{fl}{sp}"] = rs.apply_jacobian(U, R, v)[:, :4]
            x, it = rs.solve_gmres(U, R, -R)
            out[f"gmres_x_{fl}{sp}"] = x[:, :4]
            out[f"gmres_it_{fl}{sp}"] = it
    out["R_schemes"] = np.stack(R_all)
    out["schemes"] = np.array(SCHEMES, np.int32)
    for fl, sp, steps in ((0, 0, 20), (2, 2, 20)):
        rs = oracle.RefSolver(rm, 0.5, 101325.0, 300.0, fl, sp)
        out[f"newton_{fl}{sp}_steps"] = steps
        out[f"newton_{fl}{sp}_U"] = rs.solve_implicit(U0, 1e-12, steps)[:, :4]
    rs = oracle.RefSolver(rm, 0.5, 101325.0, 300.0, 2, 2)
    out["explicit_U"] = rs.time_step_explicit(U, 1e-6)[:, :4]
    np.savez_compressed(os.path.join(OUT, "ref_bump_1230_th.npz"), **out)


def gen_wedge(tmp: str):
    m = cfd.Mesh.generate(1, 0.25)
    path = os.path.join(tmp, "wedge_q.msh")
    m.write_gmsh(path)
    rm = oracle.RefMesh(path)
    out = mesh_arrays(path, rm)
    from mycode_amd import configs
    p = configs.params_for(1)
    U = configs.initial_state(1, m, p)
    out.update(U=U)
    for fl, sp in ((2, 2), (0, 0)):
        rs = oracle.RefSolver(rm, 2.5, 101325.0, 300.0, fl, sp)
        out[f"R_{fl}{sp}"] = rs.compute_residuals(U)[:, :4]
    np.savez_compressed(os.path.join(OUT, "ref_wedge_q.npz"), **out)


def gen_next(tmp: str):
    import hashlib
    src = os.path.join(oracle.REFERENCE, "mesh", "bump_1230_th")
    path = lf_copy(src, tmp)
    rm = oracle.RefMesh(path)
    rs = oracle.RefSolver(rm, 0.5, 101325.0, 300.0, 1, 1)
    U = rs.initialize()
    U[:, :4] *= 1 + 0.01 * np.random.default_rng(3).standard_normal((U.shape[0], 4))
    out = {"U_in": U.copy(), "dt_phys": np.float64(1e-5), "n_steps": np.int32(2)}
    Uu = rs.solve_unsteady(U, 1e-5, 2)
    out["unsteady_U"] = Uu
    out["entropy_printed"] = np.float64(rs.entropy_error(Uu))
    x_target = float(rm.cells()["x_c"][100])
    csv = os.path.join(tmp, "bl.csv")
    rs.bl_profile_csv(Uu, x_target, csv)
    out["bl_x"] = np.float64(x_target)
    out["bl_rows"] = np.loadtxt(csv, delimiter=",", skiprows=1, ndmin=2)
    os.makedirs(os.path.join(tmp, "results"), exist_ok=True)
    rm.save_vtk(Uu, 7)
    data = open(os.path.join(tmp, "results", "sol_7.vtk"), "rb").read()
    out["vtk_sha256"] = np.array(hashlib.sha256(data).hexdigest())
    out["vtk_bytes"] = np.int64(len(data))
    np.savez_compressed(os.path.join(OUT, "ref_bump_1230_next.npz"), **out)


def main():
    if not oracle.ref_available():
        oracle.build(ref=True)
    with tempfile.TemporaryDirectory() as tmp:
        cwd = os.getcwd()
        os.chdir(tmp)  # the reference writes results/residual_history.csv only if results/ exists
        try:
            gen_bump(tmp)
            gen_wedge(tmp)
            gen_next(tmp)
        finally:
            os.chdir(cwd)
    for f in sorted(os.listdir(OUT)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(OUT, f)))


if __name__ == "__main__":
    main()
