"""The FP64 operation counter of the CPU restatement (the op-counting scalar
behind bench.py's roofline_fp64): running the residual on it changes no
value (bit-identical to the plain residual on every configuration's physics),
and the count is a property of the discretisation, not of the mesh size: per
cell it agrees within 1 % between a mesh and one with four times the cells."""
import numpy as np
import pytest

from mycode_amd import configs
from oracle import oracle


def _solver(cfg, scale):
    mesh, p, U = configs.setup(cfg, scale)
    U = U.copy()
    U[:, :4] *= 1.0 + 1e-3 * np.random.default_rng(cfg).standard_normal((U.shape[0], 4))
    return mesh, oracle.OracleSolver(oracle.OracleMesh.from_mesh(mesh), p), U


@pytest.mark.parametrize("cfg,scale", [(1, 0.05), (2, 0.1), (3, 0.05), (4, 0.1), (5, 0.02)])
def test_counted_residual_is_the_residual(cfg, scale):
    mesh, o, U = _solver(cfg, scale)
    tot, cell, R = o.residual_opcount(U)
    assert np.array_equal(R, o.compute_residuals(U))
    assert tot["flops"] == tot["add"] + tot["mul"] + tot["div"] + tot["sqrt"]
    assert 0 < cell["flops"] < tot["flops"]
    for k in oracle.OracleSolver.OPS:
        assert cell[k] <= tot[k]
    print(cfg, mesh.n_cells, tot, f"{tot['flops'] / mesh.n_cells:.1f} flops per cell")


@pytest.mark.parametrize("cfg,scales", [(4, (0.1, 0.2)), (3, (0.05, 0.1))])
def test_count_per_cell_is_mesh_independent(cfg, scales):
    per = []
    for sc in scales:
        mesh, o, U = _solver(cfg, sc)
        tot, cell, _ = o.residual_opcount(U)
        per.append((tot["flops"] / mesh.n_cells, cell["flops"] / mesh.n_cells))
    assert abs(per[1][0] - per[0][0]) <= 0.01 * per[1][0]
    assert abs(per[1][1] - per[0][1]) <= 0.01 * per[1][1]
