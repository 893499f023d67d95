"""Make sst_plate_cf_ref.npz: the wall skin friction of a tightly converged
FP64 solution of the turbulent flat plate (configuration 3, 1,320,146
cells), the reference the bench's JFNK legs are measured against.

The solve: tangent Jv with FP64 matvec, preconditioner (line-implicit block
Gauss-Seidel) and Krylov basis, GMRES(60) with two restarts at a 1e-3
tolerance, the adaptive CFL controller, to a 1e-13 drop of the
volume-weighted residual norm (or 1200 Newton steps).  Stored: Cf on the
no-slip faces in reference face order and their x_mid, the drop reached
(summed and per equation), the Newton steps and the mesh hash.

Usage (on a GPU):  python mycode_amd/data/make_sst_plate_ref.py [OUT.npz]
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from mycode_amd import cfd, configs  # noqa: E402


def main(out):
    mesh, p, U0 = configs.setup(3, 1.0)
    s = cfd.Solver(mesh, p)
    g = s.gmres_opts(m=60, tol=1e-3, jv=1, precision=64, precond=3, restarts=2)
    o = s.newton_opts(gmres=g, tolerance=0.0, rel_drop=1e-13, max_steps=1200, controller=2, cfl0=1.0,
                      cfl_max=1e12, log=25)
    t = time.time()
    U, res, hist = s.solve_implicit(U0, o)
    _, Cf, _ = s.wall_outputs(U)
    vol = mesh.cells()["volume"]

    def vm(R):
        return np.sqrt(np.sum(R[:, :6] ** 2 / vol[:, None], 0) / np.sum(vol))

    d = vm(s.compute_residuals(U)) / vm(s.compute_residuals(U0))
    f = mesh.faces()
    w = (f["neighbor"] < 0) & (f["bc_tag"] == 4)
    np.savez_compressed(out, Cf=Cf[w], x_cf=f["x_mid"][w], drop=hist["vm"][-1] / hist["vm"][0],
                        drop_per_equation=d, steps=res.steps, mesh_hash=np.uint64(mesh.hash()))
    print(json.dumps({"out": out, "steps": res.steps, "converged": bool(res.converged),
                      "drop": float(hist["vm"][-1] / hist["vm"][0]), "drop_per_equation": [float(x) for x in d],
                      "seconds": time.time() - t}), flush=True)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else os.path.join(HERE, "sst_plate_cf_ref.npz"))
