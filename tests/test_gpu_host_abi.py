"""A compiled C++ host (tests/host_abi_test.cpp, built by
__graft_entry__.build()) passes a std::vector of the reference's 48-byte
State through cfd_residual, cfd_local_timestep and cfd_jv_fd using only
include/cfd_gpu.h.  Its outputs equal, bit for bit, what the Python binding
gets from the same library for the same state."""
import os
import subprocess

import numpy as np
import pytest

from mycode_amd import cfd

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "host_abi_test")


def test_cpp_host_through_the_c_abi(tmp_path):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} is missing: run __graft_entry__.build() first")
    r = subprocess.run([BIN, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    n = int(r.stdout.strip().split("N=")[1])
    U = np.fromfile(tmp_path / "U.bin").reshape(n, 6)
    R = np.fromfile(tmp_path / "R.bin").reshape(n, 6)
    Jv = np.fromfile(tmp_path / "Jv.bin").reshape(n, 6)
    mesh = cfd.Mesh.generate(2, 0.1)
    assert mesh.n_cells == n
    p = cfd.default_params()
    cfd.flow_conditions(p, 0.2, 101325.0, 300.0)
    cfd.kinematic_viscosity(p, 1e5, 1.0)
    p.flux, p.spatial, p.bcset, p.viscous, p.nvar = cfd.ROE, cfd.SECOND_ORDER_CENTRAL, cfd.BC_PLATE, 1, 4
    s = cfd.Solver(mesh, p)
    Rp = s.compute_residuals(U, True)
    assert np.array_equal(R, Rp)
    assert np.abs(R[:, :4]).max() > 0
    s.compute_local_time_step(U, 2.0)
    v = np.zeros_like(U)
    v[:, 0], v[:, 1], v[:, 2], v[:, 3] = 1e-3 * U[:, 0], -2e-3 * U[:, 1], 3e-3 * U[:, 1], 1e-3 * U[:, 3]
    assert np.array_equal(Jv, s.apply_jacobian(U, Rp, v))
