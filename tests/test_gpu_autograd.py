"""The forward tangent and the discrete adjoint as a torch.autograd.Function.

`_residual_function` wraps the HIP residual R(U) on device tensors (N x 6,
the reference's AoS order): its forward-mode rule is the forward tangent
(cfd_dev_jv_tangent) and its backward rule the reverse product
(cfd_dev_vjp_adjoint), both through the C-ABI.  Both library products carry
the pseudo-time term V/dt of the Newton system (Solver.cpp:1381-1384); the
rules evaluate it at CFL 1e12 and take it out, so they are dR/dU v and
dR/dU^T w.  The limiter is frozen at the forward state, as in the
reference's FD Jv (Solver.cpp:1366).

Pinned by: the oracle's complex-step derivative (independent CPU arithmetic,
h = 1e-30), central differences of the HIP residual itself, and the
reference's FD Jv (cfd_jv_fd, Solver.cpp:1331-1386).
"""
import numpy as np
import pytest
import torch
from torch.autograd import forward_ad as fwAD

from mycode_amd import cfd, configs
from oracle import oracle

pytestmark = pytest.mark.gpu

CASES = [(1, 0.1), (2, 0.05), (3, 0.02), (4, 0.03)]
CFL_FREEZE = 1e12  # V/dt = lambda / CFL: the time term, taken out again, is ~1e-12 of the flux terms


def _residual_function(s, nv):
    dev = torch.device("cuda:0")
    N, p = s.N, s.pitch
    vol = torch.from_numpy(s.mesh.cells()["volume"]).to(dev)

    def soa(x6):  # (N, 6) in reference order -> nv x pitch SoA on the device
        out = torch.zeros(nv * p, dtype=torch.float64, device=dev)
        x6 = x6.detach().contiguous()
        torch.cuda.synchronize()  # torch's stream has written x6 and out
        s.dev_pack(x6.data_ptr(), out.data_ptr())
        s.synchronize()
        return out

    def aos(xd):
        out = torch.zeros(N, 6, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        s.dev_unpack(xd.data_ptr(), out.data_ptr())
        s.synchronize()
        return out

    def time_term():
        dt = torch.from_numpy(s.dt_local()).to(dev)
        return (vol / (dt + 1e-30))[:, None]

    class Residual(torch.autograd.Function):
        @staticmethod
        def forward(ctx, U6):
            Ud = soa(U6)
            Rd = torch.zeros_like(Ud)
            torch.cuda.synchronize()
            s.dev_residual(Ud.data_ptr(), Rd.data_ptr(), True)
            s.dev_local_timestep(Ud.data_ptr(), CFL_FREEZE)
            s.synchronize()
            ctx.Ud = Ud
            return aos(Rd)

        @staticmethod
        def backward(ctx, g):
            w = soa(g)
            out = torch.zeros_like(w)
            torch.cuda.synchronize()
            s.dev_vjp_adjoint(ctx.Ud.data_ptr(), w.data_ptr(), out.data_ptr())
            s.synchronize()
            gU = aos(out) - time_term() * g
            gU[:, nv:] = 0
            return gU

        @staticmethod
        def jvp(ctx, v):
            vd = soa(v)
            out = torch.zeros_like(vd)
            torch.cuda.synchronize()
            s.dev_jv_tangent(ctx.Ud.data_ptr(), vd.data_ptr(), out.data_ptr(), 64)
            s.synchronize()
            Jv = aos(out) - time_term() * v
            Jv[:, nv:] = 0
            return Jv

    return Residual


def _setup(config, scale, seed=0):
    mesh, p, U = configs.setup(config, scale)
    s = cfd.Solver(mesh, p)
    o = oracle.OracleSolver(oracle.OracleMesh.from_mesh(mesh), p)
    rng = np.random.default_rng(100 * config + seed)
    U = U.copy()
    U[:, :4] *= 1 + 0.01 * rng.standard_normal((U.shape[0], 4))
    U[:, 2] += 0.01 * U[:, 1] * rng.standard_normal(U.shape[0])  # off the |u_n| = 0 kink
    if p.nvar == 6:
        U[:, 4] *= 1 + 0.5 * rng.random(U.shape[0])
        U[:, 5] *= 1 + 0.5 * rng.random(U.shape[0])
    else:
        U[:, 4:] = 0
    v = rng.standard_normal(U.shape) * np.abs(U).mean(0)
    v[:, p.nvar:] = 0
    w = rng.standard_normal(U.shape)
    w[:, p.nvar:] = 0
    return s, o, p, U, v, w


def _cs(o, U, v):
    o.compute_residuals(U, True)
    _, dR = o.residual_complex_step(U, v, 1e-30, update_limiters=False)
    return dR


@pytest.mark.parametrize("config,scale", CASES)
def test_forward_mode_vs_complex_step(config, scale):
    s, o, p, U, v, w = _setup(config, scale)
    print(f"config={config} N={s.N}", flush=True)
    Res = _residual_function(s, p.nvar)
    dev = torch.device("cuda:0")
    Ut, vt = torch.from_numpy(U).to(dev), torch.from_numpy(v).to(dev)
    with fwAD.dual_level():
        out = Res.apply(fwAD.make_dual(Ut, vt))
        R, Jv = fwAD.unpack_dual(out)
    Jv = Jv.cpu().numpy()
    ref = _cs(o, U, v)
    assert np.linalg.norm(Jv - ref) <= 1e-10 * np.linalg.norm(ref)
    # the primal is the residual of the host path bit for bit
    assert np.array_equal(R.cpu().numpy(), s.compute_residuals(U, True))


@pytest.mark.parametrize("config,scale", CASES)
def test_backward_vs_complex_step(config, scale):
    """<dL/dU, v> with L = <w, R(U)> by backward, against <w, dR v> by complex step."""
    s, o, p, U, v, w = _setup(config, scale, seed=1)
    print(f"config={config} N={s.N}", flush=True)
    Res = _residual_function(s, p.nvar)
    dev = torch.device("cuda:0")
    Ut = torch.from_numpy(U).to(dev).requires_grad_(True)
    L = (Res.apply(Ut) * torch.from_numpy(w).to(dev)).sum()
    (g,) = torch.autograd.grad(L, Ut)
    lhs = float((g.cpu().numpy() * v).sum())
    dR = _cs(o, U, v)
    rhs = float((w * dR).sum())
    assert abs(lhs - rhs) <= 1e-11 * np.linalg.norm(w) * np.linalg.norm(dR)


@pytest.mark.parametrize("config,scale", [(3, 0.02), (4, 0.03)])
def test_forward_mode_vs_central_differences(config, scale):
    """Second-order central schemes without a limiter: R is smooth along v away
    from the switches, so (R(U + h v) - R(U - h v)) / 2h matches to O(h^2)."""
    s, o, p, U, v, w = _setup(config, scale, seed=2)
    Res = _residual_function(s, p.nvar)
    dev = torch.device("cuda:0")
    Ut, vt = torch.from_numpy(U).to(dev), torch.from_numpy(v).to(dev)
    with fwAD.dual_level():
        Jv = fwAD.unpack_dual(Res.apply(fwAD.make_dual(Ut, vt))).tangent.cpu().numpy()
    h = 1e-6
    Rp = Res.apply(Ut + h * vt).cpu().numpy()
    Rm = Res.apply(Ut - h * vt).cpu().numpy()
    fd = (Rp - Rm) / (2 * h)
    assert np.linalg.norm(Jv - fd) <= 1e-6 * np.linalg.norm(Jv)


@pytest.mark.parametrize("config,scale", [(2, 0.05), (4, 0.03)])
def test_backward_vs_reference_fd_jv(config, scale):
    """<dL/dU, v> against <w, J_FD v> of the reference's FD product, within its
    truncation and round-off error (eps = 1e-7 (1 + |U|) / |v|)."""
    s, o, p, U, v, w = _setup(config, scale, seed=3)
    Res = _residual_function(s, p.nvar)
    dev = torch.device("cuda:0")
    Ut = torch.from_numpy(U).to(dev).requires_grad_(True)
    (g,) = torch.autograd.grad((Res.apply(Ut) * torch.from_numpy(w).to(dev)).sum(), Ut)
    h = cfd.Solver(s.mesh, p)
    R = h.compute_residuals(U, True)
    h.compute_local_time_step(U, CFL_FREEZE)
    tt = (s.mesh.cells()["volume"] / (h.dt_local() + 1e-30))[:, None]
    Jfd = h.apply_jacobian(U, R, v) - tt * v
    lhs, rhs = float((g.cpu().numpy() * v).sum()), float((w * Jfd).sum())
    assert abs(lhs - rhs) <= 1e-5 * np.linalg.norm(w) * np.linalg.norm(Jfd)
