"""Host mesh pipeline of the library: reference-exact connectivity and geometry,
RCM permutation, edge coloring, Gmsh round trip, seeded generator."""
import numpy as np
import pytest

from mycode_amd import cfd
from oracle import oracle


def test_connectivity_geometry_bitexact_vs_reference(bump_golden):
    g = bump_golden
    m = cfd.Mesh.from_arrays(g["x"], g["y"], g["cell_ptr"], g["cell_nodes"], g["bedges"])
    f, c = m.faces(), m.cells()
    for k, gk in (("owner", "f_owner"), ("neighbor", "f_neighbor"), ("bc_tag", "f_tag"), ("node1", "f_node1"),
                  ("node2", "f_node2"), ("nx", "f_nx"), ("ny", "f_ny"), ("area", "f_area"), ("x_mid", "f_xm"),
                  ("y_mid", "f_ym")):
        assert np.array_equal(f[k], g[gk]), k
    for k, gk in (("volume", "c_vol"), ("x_c", "c_xc"), ("y_c", "c_yc")):
        assert np.array_equal(c[k], g[gk]), k


@pytest.mark.parametrize("config,scale", [(1, 0.25), (2, 0.1), (3, 0.02), (4, 0.05), (5, 0.02)])
def test_generated_mesh_bitexact_vs_oracle(config, scale):
    m = cfd.Mesh.generate(config, scale)
    om = oracle.OracleMesh.from_mesh(m)
    f, of = m.faces(), om.faces()
    for k in f:
        assert np.array_equal(f[k], of[k]), k
    c, oc = m.cells(), om.cells()
    for k in c:
        assert np.array_equal(c[k], oc[k]), k


def _invariants(m):
    f, c = m.faces(), m.cells()
    N, F = m.n_cells, m.n_faces
    perm, color = m.reorder()
    # RCM permutation is a bijection
    assert np.array_equal(np.sort(perm), np.arange(N))
    # no two faces of one color share a cell
    for col in range(m.n_colors):
        sel = color == col
        cells = np.concatenate([f["owner"][sel], f["neighbor"][sel][f["neighbor"][sel] >= 0]])
        assert len(np.unique(cells)) == len(cells)
    assert m.n_colors <= 7
    # closed cells: sum of n A over each cell's faces is zero
    s = np.zeros((N, 2))
    nA = np.stack([f["nx"] * f["area"], f["ny"] * f["area"]], 1)
    np.add.at(s, f["owner"], nA)
    nb = f["neighbor"] >= 0
    np.add.at(s, f["neighbor"][nb], -nA[nb])
    assert np.max(np.abs(s)) < 1e-12 * max(1.0, f["area"].max())
    # unit normals, positive volumes
    assert np.max(np.abs(np.hypot(f["nx"], f["ny"]) - 1)) < 4e-16
    assert (c["volume"] > 0).all()
    # Euler characteristic of a simply connected planar mesh: V - E + C = 1
    assert m.n_nodes - F + N == 1
    return perm, f


@pytest.mark.parametrize("config", [1, 2, 4])
def test_rcm_coloring_invariants(config):
    m = cfd.Mesh.generate(config, 0.2)
    perm, f = _invariants(m)
    nb = f["neighbor"] >= 0
    band_new = np.max(np.abs(perm[f["owner"][nb]] - perm[f["neighbor"][nb]]))
    band_old = np.max(np.abs(f["owner"][nb] - f["neighbor"][nb]))
    assert band_new <= band_old


def test_generator_sizes_and_determinism():
    m1 = cfd.Mesh.generate(1)
    assert (m1.n_cells, m1.n_faces) == (25600, 38640)
    assert cfd.Mesh.generate(1).hash() == m1.hash()
    m2 = cfd.Mesh.generate(2)
    assert abs(m2.n_cells - 76800) < 600
    f = m1.faces()
    assert (f["neighbor"] < 0).sum() == 480


def test_gmsh_roundtrip(tmp_path):
    m = cfd.Mesh.generate(2, 0.1)
    p = str(tmp_path / "m.msh")
    m.write_gmsh(p)
    r = cfd.Mesh.read_gmsh(p)
    a, b = m.faces(), r.faces()
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    for x, y in zip(m.desc(), r.desc()):
        assert np.array_equal(x, y)


def test_mesh_errors():
    with pytest.raises(cfd.CfdError):
        cfd.Mesh.read_gmsh("/nonexistent/mesh.msh")
    # an edge shared by three cells is rejected
    x = np.array([0, 0, 1, 0, 1, -1.0])
    y = np.array([0, 0, 0, 1, 1, 0.5])
    cp = np.array([0, 3, 6, 9])
    cn = np.array([1, 2, 3, 1, 2, 4, 1, 2, 5])
    with pytest.raises(cfd.CfdError):
        cfd.Mesh.from_arrays(x, y, cp, cn, np.zeros((0, 3)))


@pytest.mark.parametrize("config,scale", [(2, 0.25), (3, 0.1), (4, 0.05)])
def test_line_segments(config, scale):
    """Line segments of the line-implicit preconditioner: a permutation of the
    cells in runs of at most 64, consecutive cells of a segment share a face,
    no two segments of one color share a face, and on the stretched
    boundary-layer meshes the wall-adjacent cells start long chains."""
    m = cfd.Mesh.generate(config, scale)
    sp, cells, cp = m.lines()
    N = m.n_cells
    assert sp[0] == 0 and sp[-1] == N and cp[0] == 0 and cp[-1] == len(sp) - 1
    assert np.array_equal(np.sort(cells), np.arange(N))
    lens = np.diff(sp)
    assert lens.min() >= 1 and lens.max() <= 64
    f = m.faces()
    inter = f["neighbor"] >= 0
    pairs = set(zip(f["owner"][inter].tolist(), f["neighbor"][inter].tolist()))
    adjacent = lambda a, b: (a, b) in pairs or (b, a) in pairs
    seg_of = np.empty(N, int)
    for k in range(len(sp) - 1):
        seg_of[cells[sp[k]:sp[k + 1]]] = k
        for t in range(sp[k] + 1, sp[k + 1]):
            assert adjacent(cells[t - 1], cells[t])
    color = np.empty(len(sp) - 1, int)
    for col in range(len(cp) - 1):
        color[cp[col]:cp[col + 1]] = col
    o, n = f["owner"][inter], f["neighbor"][inter]
    so, sn = seg_of[o], seg_of[n]
    cross = so != sn
    assert np.all(color[so[cross]] != color[sn[cross]])
    if config in (2, 3):
        assert lens.max() >= 10


@pytest.mark.parametrize("config,scale", [(1, 0.25), (2, 0.25), (3, 0.1), (4, 0.05)])
def test_cell_coloring(config, scale):
    """Cell colors of the block SGS sweeps: no two cells that share a face
    have one color, every color is used, and the iterated greedy recoloring
    keeps these meshes at four colors or fewer."""
    m = cfd.Mesh.generate(config, scale)
    nc, color = m.cell_colors()
    f = m.faces()
    inter = f["neighbor"] >= 0
    o, n = f["owner"][inter], f["neighbor"][inter]
    assert np.all(color[o] != color[n])
    assert color.min() == 0 and color.max() == nc - 1
    assert np.array_equal(np.unique(color), np.arange(nc))
    assert nc <= 4
    sp, cells, cp = m.lines()
    assert len(cp) - 1 <= 4
