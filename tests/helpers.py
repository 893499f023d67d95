"""Shared helpers of the test suite (parity metrics, parameter set-up)."""
import numpy as np

from mycode_amd import cfd


def bump_params(g, flux, spatial):
    """Parameters of the reference as shipped: bump BCs, Mach 0.5 (main.cpp:63)."""
    p = cfd.default_params()
    cfd.flow_conditions(p, float(g["M"]), float(g["P"]), float(g["T"]))
    p.flux, p.spatial, p.bcset, p.nvar, p.viscous = int(flux), int(spatial), cfd.BC_BUMP, 4, 0
    return p


def normalized_residual_error(R, Rref, scale):
    """max_i |R_i - Rref_i| / S_i with S_i = sum of |F_f| A_f over the cell's faces
    (per component, floored by the global mean), the metric that stays meaningful
    where the residual itself cancels to round-off."""
    S = np.maximum(scale, 1e-300)
    return float(np.max(np.abs(R - Rref) / S))


def flux_scale(mesh_faces, N, Fabs):
    """Sum over faces of |F A| per cell; Fabs is (F, k) of |flux * area|."""
    S = np.zeros((N, Fabs.shape[1]))
    np.add.at(S, mesh_faces["owner"], Fabs)
    nb = mesh_faces["neighbor"] >= 0
    np.add.at(S, mesh_faces["neighbor"][nb], Fabs[nb])
    return S


def max_ulp(a, b):
    """Largest distance in units in the last place between two float64 arrays."""
    a = np.ascontiguousarray(a, np.float64).view(np.int64)
    b = np.ascontiguousarray(b, np.float64).view(np.int64)
    a = np.where(a < 0, np.int64(-0x8000000000000000) - a, a)
    b = np.where(b < 0, np.int64(-0x8000000000000000) - b, b)
    return int(np.max(np.abs(a - b))) if a.size else 0


def uniform_quad_mesh(nx, ny, Lx, Ly, tags=(4, 2, 5, 1)):
    """Mesh description of nx x ny equal rectangles on [0, Lx] x [0, Ly]:
    1-based node ids (x[0], y[0] unused), CCW quads row by row, and boundary
    edges tagged (bottom, right, top, left)."""
    xs, ys = np.linspace(0.0, Lx, nx + 1), np.linspace(0.0, Ly, ny + 1)
    X, Y = np.meshgrid(xs, ys)  # (ny + 1, nx + 1), row j = y index
    x = np.concatenate([[0.0], X.ravel()])
    y = np.concatenate([[0.0], Y.ravel()])

    def nid(i, j):
        return 1 + j * (nx + 1) + i

    cells = [[nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)] for j in range(ny) for i in range(nx)]
    cell_ptr = np.arange(0, 4 * len(cells) + 1, 4, dtype=np.int32)
    cell_nodes = np.array(cells, np.int32).ravel()
    b, r, t, l = tags
    be = [(nid(i, 0), nid(i + 1, 0), b) for i in range(nx)]
    be += [(nid(nx, j), nid(nx, j + 1), r) for j in range(ny)]
    be += [(nid(i, ny), nid(i + 1, ny), t) for i in range(nx)]
    be += [(nid(0, j), nid(0, j + 1), l) for j in range(ny)]
    return x, y, cell_ptr, cell_nodes, np.array(be, np.int32)


def interior_cells(faces, n_cells, rings=1):
    """Cells at least `rings` face-neighbours away from every boundary face."""
    b = faces["neighbor"] < 0
    near = np.zeros(n_cells, bool)
    near[faces["owner"][b]] = True
    inter = ~b
    o, n = faces["owner"][inter], faces["neighbor"][inter]
    for _ in range(rings):
        nxt = near.copy()
        nxt[o[near[n]]] = True
        nxt[n[near[o]]] = True
        near = nxt
    return np.where(~near)[0]
