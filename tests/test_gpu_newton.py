"""Newton / pseudo-transient driver with the adaptive CFL controller, and the
device-pointer GMRES entry point.

Bars: the laminar plate (4,814 cells) reaches a 1e-10 drop of the
volume-weighted residual norm within 200 Newton steps with FD + FP64 Krylov
and with tangent + FP32 Krylov; their wall Cp and Cf agree within 1e-8 (the
target Cf bar is 1e-6); the device GMRES gives the host GMRES's
iteration count and solution to 1e-12 relative.
"""
import numpy as np
import pytest

from mycode_amd import cfd, configs

pytestmark = pytest.mark.gpu


def _solve(mesh, p, U0, jv, prec):
    s = cfd.Solver(mesh, p)
    g = s.gmres_opts(m=40, tol=1e-2, jv=jv, precision=prec, precond=2)
    o = s.newton_opts(gmres=g, tolerance=0.0, rel_drop=1e-10, max_steps=200, controller=2, cfl0=1.0, cfl_max=1e12)
    U, res, hist = s.solve_implicit(U0, o)
    return s, U, res, hist


def test_adaptive_controller_fd64_and_tangent32_converge():
    mesh, p, U0 = configs.setup(2, 0.25)
    out = {}
    for jv, prec in ((0, 64), (1, 32)):
        s, U, res, hist = _solve(mesh, p, U0, jv, prec)
        print(f"jv={jv} FP{prec}: steps={res.steps} converged={res.converged} "
              f"drop={hist['vm'][-1] / hist['vm'][0]:.2e} gmres={res.total_gmres_iters} "
              f"cfl_end={hist['cfl'][-1]:.3g}", flush=True)
        assert res.converged and hist["vm"][-1] <= 1e-10 * hist["vm"][0]
        assert np.all(hist["cfl"] >= 0.01 * 1.0 - 1e-15) and np.all(hist["cfl"] <= 1e12)
        out[(jv, prec)] = s.wall_outputs(U)
    Cp64, Cf64, _ = out[(0, 64)]
    Cp32, Cf32, _ = out[(1, 32)]
    print(f"max |dCp| {np.max(np.abs(Cp64 - Cp32)):.2e}  max |dCf| {np.max(np.abs(Cf64 - Cf32)):.2e}")
    assert np.max(np.abs(Cp64 - Cp32)) <= 1e-8
    assert np.max(np.abs(Cf64 - Cf32)) <= 1e-8


def test_dev_gmres_matches_host_gmres():
    import torch
    mesh, p, U = configs.setup(4, 0.05)
    U = U.copy()
    U[:, :4] *= 1 + 0.01 * np.random.default_rng(8).standard_normal((U.shape[0], 4))
    s = cfd.Solver(mesh, p)
    R = s.compute_residuals(U, True)
    s.compute_local_time_step(U, 20.0)
    o = s.gmres_opts(m=40, tol=1e-3, jv=1, precision=64, precond=2)
    x_host, it_host = s.solve_gmres(U, R, -R, o)
    dev = torch.device("cuda:0")
    nv, pitch = p.nvar, s.pitch
    U6 = torch.from_numpy(np.ascontiguousarray(U)).to(dev)
    R6 = torch.from_numpy(np.ascontiguousarray(R)).to(dev)
    Ud = torch.zeros(nv * pitch, dtype=torch.float64, device=dev)
    Rd = torch.zeros_like(Ud)
    torch.cuda.synchronize()
    s.dev_pack(U6.data_ptr(), Ud.data_ptr())
    s.dev_pack(R6.data_ptr(), Rd.data_ptr())
    s.synchronize()
    rhs = -Rd
    x = torch.zeros_like(Rd)
    torch.cuda.synchronize()
    it = s.dev_gmres(Ud.data_ptr(), Rd.data_ptr(), rhs.data_ptr(), x.data_ptr(), o)
    x6 = torch.zeros(s.N * 6, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()  # torch's fill runs on its own stream, the library on the context stream
    s.dev_unpack(x.data_ptr(), x6.data_ptr())
    s.synchronize()
    xd = x6.cpu().numpy().reshape(-1, 6)
    print(f"iterations host {it_host} device {it}")
    assert it == it_host
    assert np.linalg.norm(xd - x_host) <= 1e-12 * np.linalg.norm(x_host)
