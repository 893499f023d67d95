"""Caller streams and argument checks of the C-ABI on the device.

cfd_ctx_set_stream puts the context's launches on the caller's stream, so
torch work queued before and after a cfd_dev_* call on that stream is
ordered with it without device-wide synchronisation.  The test runs the
4.4M-cell configuration-5 mesh through pack -> residual -> unpack with every
buffer allocated and filled by torch on its current stream and no
synchronisation in between; the result must be the host-pointer path's
residual bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

from mycode_amd import cfd, configs

pytestmark = pytest.mark.gpu


def test_caller_stream_orders_torch_work():
    import torch

    mesh, p, U0 = configs.setup(5, 1.0)
    print(f"caller stream: configuration 5, {mesh.n_cells} cells", flush=True)
    s = cfd.Solver(mesh, p)
    s.set_face_flux(configs.shock_band(mesh))
    U0 = U0.copy()
    U0[:, :4] *= 1 + 1e-3 * np.random.default_rng(5).standard_normal((U0.shape[0], 4))
    R_host = s.compute_residuals(U0)
    st = torch.cuda.current_stream()
    s.set_stream(st.cuda_stream)
    assert s.stream == (st.cuda_stream or 0)
    nv, P = p.nvar, s.pitch
    U6 = torch.from_numpy(U0).to("cuda", non_blocking=False)
    for rep in range(2):
        Usoa = torch.full((nv * P,), float("nan"), dtype=torch.float64, device="cuda")
        Rsoa = torch.full((nv * P,), float("nan"), dtype=torch.float64, device="cuda")
        R6 = torch.full((mesh.n_cells, 6), float("nan"), dtype=torch.float64, device="cuda")
        s.dev_pack(U6.data_ptr(), Usoa.data_ptr())
        s.dev_residual(Usoa.data_ptr(), Rsoa.data_ptr(), True)
        s.dev_unpack(Rsoa.data_ptr(), R6.data_ptr())
        got = R6.cpu().numpy()  # torch's copy waits on its stream, which now holds the library's work
        assert np.array_equal(got[:, :4], R_host[:, :4]), rep
    s.set_stream(None)
    assert s.stream != (st.cuda_stream or 0)


def test_unsteady_rejects_unsized_history():
    mesh, p, U0 = configs.setup(2, 0.05)
    s = cfd.Solver(mesh, p)
    hvm = np.zeros(4)
    hn = np.zeros(2, np.int32)
    U = np.ascontiguousarray(U0, np.float64).copy()
    rc = s.lib.cfd_solve_unsteady(s._h, U.ctypes.data_as(C.POINTER(C.c_double)), 1e-5, 2, None, 0, 1e-6, 0.7,
                                  hvm.ctypes.data_as(C.POINTER(C.c_double)), hn.ctypes.data_as(C.POINTER(C.c_int)))
    assert rc == -1
    # the binding resolves the reference's default before sizing the history
    _, h, n = s.solve_unsteady(U0, 1e-5, 1, max_newton=0)
    assert h.shape == (1, 15)


def test_products_need_a_formed_limiter():
    mesh, p, U0 = configs.setup(1, 0.1)
    p.spatial = cfd.SECOND_ORDER_LIMITED
    s = cfd.Solver(mesh, p)
    import torch

    U = torch.from_numpy(U0).cuda()
    Us = torch.zeros(p.nvar * s.pitch, dtype=torch.float64, device="cuda")
    v = torch.ones_like(Us)
    Jv = torch.zeros_like(Us)
    torch.cuda.synchronize()
    s.dev_pack(U.data_ptr(), Us.data_ptr())
    s.dev_local_timestep(Us.data_ptr(), 1.0)
    with pytest.raises(cfd.CfdError):
        s.dev_jv_tangent(Us.data_ptr(), v.data_ptr(), Jv.data_ptr(), 64)
    with pytest.raises(cfd.CfdError):
        s.dev_vjp_adjoint(Us.data_ptr(), v.data_ptr(), Jv.data_ptr())
    s.dev_residual(Us.data_ptr(), Jv.data_ptr(), True)
    s.dev_jv_tangent(Us.data_ptr(), v.data_ptr(), Jv.data_ptr(), 64)
    s.synchronize()
