"""cfd_dev_norm_vw_noalloc, the entry point that allocates nothing, against
cfd_norm_vw (Solver::norm_L2_VolumeWeighted) on two calls in a row, the
second on a smaller mesh (its static scratch still holds the first call's
partials).  Both form the same sums in the same order, so the bar is
bit-for-bit equality."""
import numpy as np
import pytest

from mycode_amd import cfd, configs

pytestmark = pytest.mark.gpu


def _case(scale, seed):
    mesh, p, U0 = configs.setup(2, scale)
    U = U0.copy()
    U[:, :4] *= 1 + 1e-2 * np.random.default_rng(seed).standard_normal((U.shape[0], 4))
    s = cfd.Solver(mesh, p)
    R = s.compute_residuals(U)
    return s, p, R


def test_noalloc_norm_matches_cfd_norm_vw():
    import torch

    cases = [_case(1.0, 1), _case(0.1, 2)]
    assert cases[1][0].N < cases[0][0].N
    for s, p, R in cases:
        print(f"noalloc norm: {s.N} cells", flush=True)
        ref = s.norm_L2_volume_weighted(R)
        st = torch.cuda.current_stream()
        s.set_stream(st.cuda_stream)
        R6 = torch.from_numpy(np.ascontiguousarray(R)).cuda()
        Rs = torch.zeros(p.nvar * s.pitch, dtype=torch.float64, device="cuda")
        out = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
        s.dev_pack(R6.data_ptr(), Rs.data_ptr())
        s.dev_norm_vw_noalloc(Rs.data_ptr(), out.data_ptr())
        got = float(out.cpu()[0])
        s.set_stream(None)
        assert ref > 0 and got == ref, (s.N, got, ref)


def test_noalloc_rejects_null_arguments():
    s, p, R = _case(0.05, 3)
    with pytest.raises(cfd.CfdError):
        s.dev_norm_vw_noalloc(0, 0)
