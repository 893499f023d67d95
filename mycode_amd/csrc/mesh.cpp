// Mesh pipeline on the host: Gmsh v2.2 reader, reference-order connectivity,
// geometry, RCM reordering, edge coloring, and the seeded synthetic meshes of
// the benchmark configurations.
//
// Connectivity and geometry reproduce Mesh::buildConnectivity
// (Mesh.cpp:102-143) and Mesh::computeGeometry (Mesh.cpp:148-202) exactly —
// same face order, owner/neighbor, tags and the same double arithmetic — but
// with an open-addressing hash table instead of std::map so the 8.4M-face
// meshes build in well under a second.
#include "mesh.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <new>
#include <numeric>
#include <sstream>
#include <string>

namespace {

// 64-bit key of an undirected edge (sorted node ids)
inline uint64_t edge_key(int a, int b) {
  uint32_t lo = (uint32_t)std::min(a, b), hi = (uint32_t)std::max(a, b);
  return ((uint64_t)lo << 32) | hi;
}

// Open-addressing map uint64 -> int (insert / find), linear probing.
struct EdgeMap {
  std::vector<uint64_t> keys;
  std::vector<int> vals;
  uint64_t mask = 0;
  static constexpr uint64_t EMPTY = ~0ull;
  explicit EdgeMap(size_t expected) {
    size_t cap = 16;
    while (cap < 2 * expected + 16) cap <<= 1;
    keys.assign(cap, EMPTY);
    vals.assign(cap, -1);
    mask = cap - 1;
  }
  static uint64_t mix(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    return k;
  }
  // returns pointer to the value slot; *fresh tells whether it was inserted
  int* slot(uint64_t k, bool* fresh) {
    uint64_t h = mix(k) & mask;
    while (true) {
      if (keys[h] == k) { *fresh = false; return &vals[h]; }
      if (keys[h] == EMPTY) { keys[h] = k; *fresh = true; return &vals[h]; }
      h = (h + 1) & mask;
    }
  }
  const int* find(uint64_t k) const {
    uint64_t h = mix(k) & mask;
    while (true) {
      if (keys[h] == k) return &vals[h];
      if (keys[h] == EMPTY) return nullptr;
      h = (h + 1) & mask;
    }
  }
};

int build_connectivity(cfd_mesh& m, std::string& err) {
  int N = m.ncells();
  size_t nedges = m.cell_nodes.size();
  EdgeMap bmap(m.bedges.size() / 3 + 1);
  for (size_t e = 0; e + 2 < m.bedges.size(); e += 3) {
    bool fresh;
    *bmap.slot(edge_key(m.bedges[e], m.bedges[e + 1]), &fresh) = m.bedges[e + 2];  // last tag wins, like map[]=
  }
  EdgeMap fmap(nedges);
  m.n1.clear(); m.n2.clear(); m.own.clear(); m.nei.clear(); m.tag.clear();
  m.n1.reserve(nedges); m.n2.reserve(nedges); m.own.reserve(nedges); m.nei.reserve(nedges); m.tag.reserve(nedges);
  for (int c = 0; c < N; ++c) {
    int b = m.cell_ptr[c], n = m.cell_ptr[c + 1] - b;
    if (n != 3 && n != 4) { err = "cell with " + std::to_string(n) + " nodes"; return CFD_ERR_MESH; }
    for (int j = 0; j < n; ++j) {
      int a1 = m.cell_nodes[b + j], a2 = m.cell_nodes[b + (j + 1) % n];
      if (a1 < 1 || a1 > m.n_nodes || a2 < 1 || a2 > m.n_nodes) { err = "node id out of range"; return CFD_ERR_MESH; }
      bool fresh;
      int* v = fmap.slot(edge_key(a1, a2), &fresh);
      if (!fresh) {
        int f = *v;
        if (m.nei[f] >= 0) { err = "edge shared by more than two cells"; return CFD_ERR_MESH; }
        m.nei[f] = c;
      } else {
        int f = (int)m.own.size();
        *v = f;
        m.n1.push_back(std::min(a1, a2));
        m.n2.push_back(std::max(a1, a2));
        m.own.push_back(c);
        m.nei.push_back(-1);
        const int* t = bmap.find(edge_key(a1, a2));
        m.tag.push_back(t ? *t : 0);
      }
    }
  }
  return CFD_OK;
}

void compute_geometry(cfd_mesh& m) {
  int N = m.ncells(), F = m.nfaces();
  const double* X = m.x.data();
  const double* Y = m.y.data();
  m.vol.assign(N, 0.0); m.xc.assign(N, 0.0); m.yc.assign(N, 0.0);
  for (int c = 0; c < N; ++c) {
    const int* nd = &m.cell_nodes[m.cell_ptr[c]];
    int n = m.cell_ptr[c + 1] - m.cell_ptr[c];
    if (n == 3) {
      double x1 = X[nd[0]], y1 = Y[nd[0]], x2 = X[nd[1]], y2 = Y[nd[1]], x3 = X[nd[2]], y3 = Y[nd[2]];
      m.vol[c] = 0.5 * std::abs((x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1));
      m.xc[c] = (x1 + x2 + x3) / 3.0;
      m.yc[c] = (y1 + y2 + y3) / 3.0;
    } else {
      double sx = 0, sy = 0;
      for (int k = 0; k < 4; ++k) { sx += X[nd[k]]; sy += Y[nd[k]]; }
      m.xc[c] = sx / 4.0;
      m.yc[c] = sy / 4.0;
      double a = 0.0;
      for (int k = 0; k < 4; ++k) {
        int p = nd[k], q = nd[(k + 1) % 4];
        a += (X[p] * Y[q] - X[q] * Y[p]);
      }
      m.vol[c] = 0.5 * std::abs(a);
    }
  }
  m.nx.assign(F, 0); m.ny.assign(F, 0); m.area.assign(F, 0); m.xm.assign(F, 0); m.ym.assign(F, 0);
  for (int f = 0; f < F; ++f) {
    double dx = X[m.n2[f]] - X[m.n1[f]];
    double dy = Y[m.n2[f]] - Y[m.n1[f]];
    double A = std::sqrt(dx * dx + dy * dy);
    m.area[f] = A;
    m.xm[f] = 0.5 * (X[m.n1[f]] + X[m.n2[f]]);
    m.ym[f] = 0.5 * (Y[m.n1[f]] + Y[m.n2[f]]);
    double nx = dy / A, ny = -dx / A;
    double vx = m.xm[f] - m.xc[m.own[f]], vy = m.ym[f] - m.yc[m.own[f]];
    if (vx * nx + vy * ny < 0) { nx = -nx; ny = -ny; }
    m.nx[f] = nx;
    m.ny[f] = ny;
  }
}

// Reverse Cuthill-McKee over the cell graph (interior faces).  Deterministic:
// each component starts at its lowest-degree unvisited cell (lowest index on
// ties) and neighbours are queued by ascending degree, then index.
void rcm(cfd_mesh& m) {
  int N = m.ncells(), F = m.nfaces();
  std::vector<int> deg(N, 0);
  for (int f = 0; f < F; ++f)
    if (m.nei[f] >= 0) { deg[m.own[f]]++; deg[m.nei[f]]++; }
  std::vector<int> aptr(N + 1, 0);
  for (int c = 0; c < N; ++c) aptr[c + 1] = aptr[c] + deg[c];
  std::vector<int> adj(aptr[N]), fill(aptr.begin(), aptr.end() - 1);
  for (int f = 0; f < F; ++f)
    if (m.nei[f] >= 0) { adj[fill[m.own[f]]++] = m.nei[f]; adj[fill[m.nei[f]]++] = m.own[f]; }
  std::vector<int> start(N);
  std::iota(start.begin(), start.end(), 0);
  std::stable_sort(start.begin(), start.end(), [&](int a, int b) { return deg[a] < deg[b]; });
  std::vector<char> seen(N, 0);
  std::vector<int> order;
  order.reserve(N);
  size_t si = 0;
  int nb[8];
  while ((int)order.size() < N) {
    while (seen[start[si]]) ++si;
    int s = start[si];
    seen[s] = 1;
    size_t head = order.size();
    order.push_back(s);
    while (head < order.size()) {
      int c = order[head++];
      int k = 0;
      for (int a = aptr[c]; a < aptr[c + 1]; ++a)
        if (!seen[adj[a]]) { nb[k++] = adj[a]; seen[adj[a]] = 1; }
      std::sort(nb, nb + k, [&](int a, int b) { return deg[a] != deg[b] ? deg[a] < deg[b] : a < b; });
      for (int i = 0; i < k; ++i) order.push_back(nb[i]);
    }
  }
  std::reverse(order.begin(), order.end());
  m.iperm = order;
  m.perm.assign(N, 0);
  for (int i = 0; i < N; ++i) m.perm[order[i]] = i;
}

// Greedy edge coloring: faces visited by ascending RCM owner position (then
// reference index); each takes the lowest color not yet used at its owner or
// neighbor.  No two faces of one color share a cell, so a color's scatter
// needs no atomics.  Device order: by color, then owner position, then
// reference index.
void color_faces(cfd_mesh& m) {
  int N = m.ncells(), F = m.nfaces();
  std::vector<int> ord(F);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return m.perm[m.own[a]] < m.perm[m.own[b]]; });
  std::vector<uint32_t> used(N, 0);
  m.color.assign(F, 0);
  int nc = 0;
  for (int f : ord) {
    uint32_t u = used[m.own[f]] | (m.nei[f] >= 0 ? used[m.nei[f]] : 0u);
    int c = 0;
    while (u & (1u << c)) ++c;
    m.color[f] = c;
    used[m.own[f]] |= 1u << c;
    if (m.nei[f] >= 0) used[m.nei[f]] |= 1u << c;
    nc = std::max(nc, c + 1);
  }
  m.ncolors = nc;
  m.face_order = ord;
  std::stable_sort(m.face_order.begin(), m.face_order.end(), [&](int a, int b) { return m.color[a] < m.color[b]; });
  m.color_ptr.assign(nc + 1, 0);
  for (int f = 0; f < F; ++f) m.color_ptr[m.color[f] + 1]++;
  for (int c = 0; c < nc; ++c) m.color_ptr[c + 1] += m.color_ptr[c];
}

// Greedy distance-1 coloring of the cells in RCM order (each cell takes the
// lowest color not used by an already colored neighbor): the color classes
// of the multicolor Gauss-Seidel sweeps.
void color_cells(cfd_mesh& m) {
  int N = m.ncells(), F = m.nfaces();
  std::vector<int> deg(N + 1, 0);
  for (int f = 0; f < F; ++f)
    if (m.nei[f] >= 0) { deg[m.perm[m.own[f]] + 1]++; deg[m.perm[m.nei[f]] + 1]++; }
  for (int i = 0; i < N; ++i) deg[i + 1] += deg[i];
  std::vector<int> adj(deg[N]), fill(deg.begin(), deg.end() - 1);
  for (int f = 0; f < F; ++f)
    if (m.nei[f] >= 0) {
      int a = m.perm[m.own[f]], b = m.perm[m.nei[f]];
      adj[fill[a]++] = b;
      adj[fill[b]++] = a;
    }
  m.cell_color.assign(N, -1);
  int nc = 0;
  for (int i = 0; i < N; ++i) {
    uint32_t used = 0;
    for (int k = deg[i]; k < deg[i + 1]; ++k)
      if (m.cell_color[adj[k]] >= 0) used |= 1u << m.cell_color[adj[k]];
    int c = 0;
    while (used & (1u << c)) ++c;
    m.cell_color[i] = c;
    nc = std::max(nc, c + 1);
  }
  // iterated greedy, as for the line segments below: whole classes in reverse
  // order, which never adds a color and usually folds the sparse last ones
  for (int round = 0; round < 16 && nc > 2; ++round) {
    std::vector<int> ncol(N, -1);
    int nn = 0;
    for (int col = nc - 1; col >= 0; --col)
      for (int i = 0; i < N; ++i) {
        if (m.cell_color[i] != col) continue;
        uint32_t used = 0;
        for (int k = deg[i]; k < deg[i + 1]; ++k)
          if (ncol[adj[k]] >= 0) used |= 1u << ncol[adj[k]];
        int c = 0;
        while (used & (1u << c)) ++c;
        ncol[i] = c;
        nn = std::max(nn, c + 1);
      }
    m.cell_color.swap(ncol);
    nc = nn;
  }
  m.ccell_ptr.assign(nc + 1, 0);
  for (int i = 0; i < N; ++i) m.ccell_ptr[m.cell_color[i] + 1]++;
  for (int c = 0; c < nc; ++c) m.ccell_ptr[c + 1] += m.ccell_ptr[c];
  std::vector<int> pos(m.ccell_ptr.begin(), m.ccell_ptr.end() - 1);
  m.ccells.assign(N, 0);
  for (int i = 0; i < N; ++i) m.ccells[pos[m.cell_color[i]]++] = i;
}

// Line segments for the line-implicit preconditioner.  The coupling of two
// cells through face f is w = A_f / |(x_n - x_o) . n_f| (the two-point
// diffusion coefficient); a cell is anisotropic when its largest coupling is
// at least 4 times its smallest.  From the most anisotropic unassigned cells,
// chains grow in both directions through each cell's strongest coupling
// (other than the one it came from) while the next cell is unassigned and
// anisotropic and that face is at least half of both cells' strongest
// couplings.  Chains are cut into segments of at most 64 cells; every other
// cell is a segment of its own.  Segments are colored greedily so that no two
// adjacent segments share a color.
void build_lines(cfd_mesh& m) {
  const int N = m.ncells(), F = m.nfaces();
  constexpr int LMAX = 64;
  // per RCM position: up to 4 slots in ascending reference face index
  std::vector<int> sl_nb((size_t)4 * N, -1), cnt(N, 0);
  std::vector<double> sl_w((size_t)4 * N, 0.0);
  for (int f = 0; f < F; ++f) {
    const int o = m.perm[m.own[f]];
    const int n = m.nei[f] >= 0 ? m.perm[m.nei[f]] : -1;
    double w = 0.0;
    if (n >= 0) {
      const int ro = m.own[f], rn = m.nei[f];
      const double d = std::abs((m.xc[rn] - m.xc[ro]) * m.nx[f] + (m.yc[rn] - m.yc[ro]) * m.ny[f]);
      w = m.area[f] / std::max(d, 1e-300);
    }
    sl_nb[(size_t)4 * o + cnt[o]] = n;
    sl_w[(size_t)4 * o + cnt[o]++] = w;
    if (n >= 0) {
      sl_nb[(size_t)4 * n + cnt[n]] = o;
      sl_w[(size_t)4 * n + cnt[n]++] = w;
    }
  }
  std::vector<double> wmax(N, 0.0), ratio(N, 1.0);
  for (int i = 0; i < N; ++i) {
    double lo = 1e300, hi = 0.0;
    int ni = 0;
    for (int q = 0; q < 4; ++q)
      if (sl_nb[(size_t)4 * i + q] >= 0) {
        lo = std::min(lo, sl_w[(size_t)4 * i + q]);
        hi = std::max(hi, sl_w[(size_t)4 * i + q]);
        ++ni;
      }
    wmax[i] = hi;
    ratio[i] = ni >= 2 ? hi / lo : 1.0;
  }
  auto aniso = [&](int i) { return ratio[i] >= 4.0; };
  auto link = [&](int a, int b) {  // coupling of a to its neighbor b
    for (int q = 0; q < 4; ++q)
      if (sl_nb[(size_t)4 * a + q] == b) return sl_w[(size_t)4 * a + q];
    return 0.0;
  };
  // the strongest neighbor of `cur` other than `avoid`, if the chain may continue there
  std::vector<char> used(N, 0);
  std::vector<int> chain_id(N, -1);
  int cur_chain = 0;
  auto step = [&](int cur, int avoid) {
    int best = -1;
    double bw = 0.0;
    for (int q = 0; q < 4; ++q) {
      const int b = sl_nb[(size_t)4 * cur + q];
      if (b < 0 || b == avoid) continue;
      if (sl_w[(size_t)4 * cur + q] > bw) {
        bw = sl_w[(size_t)4 * cur + q];
        best = b;
      }
    }
    if (best < 0 || used[best] || !aniso(best)) return -1;
    if (bw < 0.5 * wmax[cur] || link(best, cur) < 0.5 * wmax[best]) return -1;
    for (int q = 0; q < 4; ++q) {  // a chain touches itself only between consecutive cells
      const int b = sl_nb[(size_t)4 * best + q];
      if (b >= 0 && b != cur && chain_id[b] == cur_chain) return -1;
    }
    return best;
  };
  std::vector<int> order(N);
  for (int i = 0; i < N; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return ratio[a] > ratio[b]; });
  std::vector<std::vector<int>> segs;
  for (int s : order) {
    if (used[s]) continue;
    std::vector<int> chain{s};
    used[s] = 1;
    chain_id[s] = ++cur_chain;
    if (aniso(s)) {
      for (int cur = s, prev = -1;;) {  // forward
        const int nx = step(cur, prev);
        if (nx < 0) break;
        chain.push_back(nx);
        used[nx] = 1;
        chain_id[nx] = cur_chain;
        prev = cur;
        cur = nx;
      }
      std::vector<int> back;
      for (int cur = s, prev = chain.size() > 1 ? chain[1] : -1;;) {  // backward
        const int nx = step(cur, prev);
        if (nx < 0) break;
        back.push_back(nx);
        used[nx] = 1;
        chain_id[nx] = cur_chain;
        prev = cur;
        cur = nx;
      }
      chain.insert(chain.begin(), back.rbegin(), back.rend());
    }
    for (size_t b = 0; b < chain.size(); b += LMAX)
      segs.emplace_back(chain.begin() + b, chain.begin() + std::min(chain.size(), b + LMAX));
  }
  // segment coloring (segments in order of their smallest cell position)
  const int ns = (int)segs.size();
  std::vector<int> seg_of(N, -1);
  for (int k = 0; k < ns; ++k)
    for (int c : segs[k]) seg_of[c] = k;
  std::vector<int> sorder(ns);
  for (int k = 0; k < ns; ++k) sorder[k] = k;
  std::vector<int> smin(ns);
  for (int k = 0; k < ns; ++k) smin[k] = *std::min_element(segs[k].begin(), segs[k].end());
  std::stable_sort(sorder.begin(), sorder.end(), [&](int a, int b) { return smin[a] < smin[b]; });
  std::vector<int> scol(ns, -1);
  int nc = 0;
  std::vector<char> taken;
  for (int k : sorder) {
    taken.assign(nc + 1, 0);
    for (int c : segs[k])
      for (int q = 0; q < 4; ++q) {
        const int b = sl_nb[(size_t)4 * c + q];
        if (b >= 0 && seg_of[b] != k && scol[seg_of[b]] >= 0) taken[scol[seg_of[b]]] = 1;
      }
    int col = 0;
    while (taken[col]) ++col;
    scol[k] = col;
    nc = std::max(nc, col + 1);
  }
  // Iterated greedy: recolor the segments class by class, the classes in
  // reverse order (the late, sparse ones first), each segment taking the
  // lowest color free among its neighbours.  Visiting whole classes in turn
  // never needs more colors than the classes it starts from, and the last few
  // classes, which hold a handful of segments on these meshes yet cost two
  // sweep launches each, usually fold into earlier ones.
  for (int round = 0; round < 16 && nc > 2; ++round) {
    std::vector<int> rorder;
    rorder.reserve(ns);
    for (int col = nc - 1; col >= 0; --col)
      for (int k : sorder)
        if (scol[k] == col) rorder.push_back(k);
    std::vector<int> ncol(ns, -1);
    int nn = 0;
    for (int k : rorder) {
      taken.assign(nn + 1, 0);
      for (int c : segs[k])
        for (int q = 0; q < 4; ++q) {
          const int b = sl_nb[(size_t)4 * c + q];
          if (b >= 0 && seg_of[b] != k && ncol[seg_of[b]] >= 0) taken[ncol[seg_of[b]]] = 1;
        }
      int col = 0;
      while (taken[col]) ++col;
      ncol[k] = col;
      nn = std::max(nn, col + 1);
    }
    scol.swap(ncol);
    nc = nn;
  }
  // color-major segment list; within a color the multi-cell segments come
  // first, longest first, so that the 64-cell groups the device sweeps pack
  // them into occupy consecutive line positions (the order inside a color does
  // not change the sweep)
  std::vector<int> lst;
  lst.reserve(ns);
  m.segcol_ptr.assign(nc + 1, 0);
  for (int col = 0; col < nc; ++col) {
    m.segcol_ptr[col] = (int)lst.size();
    const size_t first = lst.size();
    for (int k : sorder)
      if (scol[k] == col) lst.push_back(k);
    std::stable_sort(lst.begin() + first, lst.end(),
                     [&](int a, int b) { return segs[a].size() > segs[b].size(); });
  }
  m.segcol_ptr[nc] = (int)lst.size();
  m.seg_ptr.assign(1, 0);
  m.seg_cells.clear();
  m.seg_cells.reserve(N);
  m.lslot_prev.assign(N, -1);
  m.lslot_next.assign(N, -1);
  auto slot_of = [&](int a, int b) {
    for (int q = 0; q < 4; ++q)
      if (sl_nb[(size_t)4 * a + q] == b) return q;
    return -1;
  };
  for (int k : lst) {
    const std::vector<int>& sg = segs[k];
    for (size_t i = 0; i < sg.size(); ++i) {
      m.seg_cells.push_back(sg[i]);
      if (i > 0) m.lslot_prev[sg[i]] = slot_of(sg[i], sg[i - 1]);
      if (i + 1 < sg.size()) m.lslot_next[sg[i]] = slot_of(sg[i], sg[i + 1]);
    }
    m.seg_ptr.push_back((int)m.seg_cells.size());
  }
}

// ---- synthetic meshes -------------------------------------------------------
struct SplitMix64 {  // Vigna's SplitMix64
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};
inline uint64_t stream_seed(int cfg, int purpose) { return 0x5EED0000ull + 16ull * (uint64_t)cfg + (uint64_t)purpose; }

// wall-normal spacing: dy_k = min(d1 r^k, dmax) with dmax chosen so the ny
// spacings sum to H (bisection), returned as node ordinates 0..H
std::vector<double> stretched(int ny, double H, double d1, double r) {
  auto total = [&](double dmax) {
    double s = 0, d = d1;
    for (int k = 0; k < ny; ++k) { s += std::min(d, dmax); d *= r; }
    return s;
  };
  std::vector<double> y(ny + 1, 0.0);
  if (d1 * ny >= H) {  // cannot stretch: uniform
    for (int j = 0; j <= ny; ++j) y[j] = H * j / ny;
    return y;
  }
  double lo = d1, hi = H;
  for (int it = 0; it < 200; ++it) {
    double mid = 0.5 * (lo + hi);
    if (total(mid) < H) lo = mid; else hi = mid;
  }
  double d = d1, s = 0;
  for (int k = 0; k < ny; ++k) { s += std::min(d, hi); y[k + 1] = s; d *= r; }
  double sc = H / y[ny];
  for (int j = 1; j < ny; ++j) y[j] *= sc;
  y[ny] = H;
  return y;
}

struct GenSpec {
  int nx, ny;
  double split_p;
};

}  // namespace

namespace cfdmesh {
int finish(cfd_mesh& m, std::string& err) {
  int rc = build_connectivity(m, err);
  if (rc) return rc;
  compute_geometry(m);
  rcm(m);
  color_faces(m);
  color_cells(m);
  build_lines(m);
  return CFD_OK;
}
}  // namespace cfdmesh

extern "C" {

int cfd_mesh_build(const cfd_mesh_desc* d, cfd_mesh** out) {
  if (!d || !out || d->n_nodes < 1 || d->n_cells < 1 || !d->x || !d->y || !d->cell_ptr || !d->cell_nodes)
    return CFD_ERR_ARG;
  std::unique_ptr<cfd_mesh> m(new (std::nothrow) cfd_mesh());
  if (!m) return CFD_ERR_ARG;
  m->n_nodes = d->n_nodes;
  m->x.assign(d->x, d->x + d->n_nodes + 1);
  m->y.assign(d->y, d->y + d->n_nodes + 1);
  m->cell_ptr.assign(d->cell_ptr, d->cell_ptr + d->n_cells + 1);
  m->cell_nodes.assign(d->cell_nodes, d->cell_nodes + d->cell_ptr[d->n_cells]);
  if (d->n_bedges > 0 && d->bedges) m->bedges.assign(d->bedges, d->bedges + 3 * d->n_bedges);
  std::string err;
  int rc = cfdmesh::finish(*m, err);
  if (rc) return rc;
  *out = m.release();
  return CFD_OK;
}

// Mesh::readGmsh (Mesh.cpp:30-97)
int cfd_mesh_read_gmsh(const char* path, cfd_mesh** out) {
  if (!path || !out) return CFD_ERR_ARG;
  std::ifstream in(path);
  if (!in.is_open()) return CFD_ERR_IO;
  std::unique_ptr<cfd_mesh> m(new cfd_mesh());
  std::string line;
  while (std::getline(in, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();  // CRLF files (Windows text mode)
    if (line == "$Nodes") {
      int nn = 0;
      in >> nn;
      m->n_nodes = nn;
      m->x.assign(nn + 1, 0.0);
      m->y.assign(nn + 1, 0.0);
      for (int i = 1; i <= nn; ++i) {
        int id;
        double z;
        in >> id >> m->x[i] >> m->y[i] >> z;
      }
      break;
    }
  }
  while (std::getline(in, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();  // CRLF files (Windows text mode)
    if (line == "$Elements") {
      int ne = 0;
      in >> ne;
      for (int e = 0; e < ne; ++e) {
        int id, type, ntags, t, phys = 0;
        in >> id >> type >> ntags;
        for (int j = 0; j < ntags; ++j) {
          in >> t;
          if (j == 0) phys = t;
        }
        if (type == 1) {
          int a, b;
          in >> a >> b;
          m->bedges.push_back(a);
          m->bedges.push_back(b);
          m->bedges.push_back(phys);
        } else if (type == 2 || type == 3) {
          int nv = type == 2 ? 3 : 4;
          for (int k = 0; k < nv; ++k) {
            int nd;
            in >> nd;
            m->cell_nodes.push_back(nd);
          }
          m->cell_ptr.push_back((int)m->cell_nodes.size());
        }
        std::getline(in, line);
      }
      break;
    }
  }
  if (m->ncells() < 1) return CFD_ERR_IO;
  std::string err;
  int rc = cfdmesh::finish(*m, err);
  if (rc) return rc;
  *out = m.release();
  return CFD_OK;
}

// New node coordinates, same topology: geometry recomputed exactly as
// Mesh::computeGeometry; face order, RCM order and colors are kept.
int cfd_mesh_set_nodes(cfd_mesh* m, const double* x, const double* y) {
  if (!m || !x || !y) return CFD_ERR_ARG;
  std::memcpy(m->x.data(), x, m->x.size() * sizeof(double));
  std::memcpy(m->y.data(), y, m->y.size() * sizeof(double));
  compute_geometry(*m);
  for (int c = 0; c < m->ncells(); ++c)
    if (!(m->vol[c] > 0)) return CFD_ERR_MESH;
  return CFD_OK;
}

int cfd_mesh_destroy(cfd_mesh* m) {
  delete m;
  return CFD_OK;
}

int cfd_mesh_sizes(const cfd_mesh* m, int* nn, int* nc, int* nf, int* ncol) {
  if (!m) return CFD_ERR_ARG;
  if (nn) *nn = m->n_nodes;
  if (nc) *nc = m->ncells();
  if (nf) *nf = m->nfaces();
  if (ncol) *ncol = m->ncolors;
  return CFD_OK;
}

int cfd_mesh_get_desc(const cfd_mesh* m, cfd_mesh_desc* d) {
  if (!m || !d) return CFD_ERR_ARG;
  d->n_nodes = m->n_nodes;
  d->x = m->x.data();
  d->y = m->y.data();
  d->n_cells = m->ncells();
  d->cell_ptr = m->cell_ptr.data();
  d->cell_nodes = m->cell_nodes.data();
  d->n_bedges = (int)m->bedges.size() / 3;
  d->bedges = m->bedges.data();
  return CFD_OK;
}

int cfd_mesh_get_faces(const cfd_mesh* m, int* n1, int* n2, int* own, int* nei, int* tag, double* nx, double* ny,
                       double* area, double* xm, double* ym) {
  if (!m) return CFD_ERR_ARG;
  size_t F = m->nfaces();
  auto cp = [F](auto* dst, const auto& src) {
    if (dst) std::memcpy(dst, src.data(), F * sizeof(src[0]));
  };
  cp(n1, m->n1); cp(n2, m->n2); cp(own, m->own); cp(nei, m->nei); cp(tag, m->tag);
  cp(nx, m->nx); cp(ny, m->ny); cp(area, m->area); cp(xm, m->xm); cp(ym, m->ym);
  return CFD_OK;
}

int cfd_mesh_get_cells(const cfd_mesh* m, double* vol, double* xc, double* yc) {
  if (!m) return CFD_ERR_ARG;
  size_t N = m->ncells();
  if (vol) std::memcpy(vol, m->vol.data(), N * 8);
  if (xc) std::memcpy(xc, m->xc.data(), N * 8);
  if (yc) std::memcpy(yc, m->yc.data(), N * 8);
  return CFD_OK;
}

int cfd_mesh_get_reorder(const cfd_mesh* m, int* perm, int* color) {
  if (!m) return CFD_ERR_ARG;
  if (perm) std::memcpy(perm, m->perm.data(), m->perm.size() * 4);
  if (color) std::memcpy(color, m->color.data(), m->color.size() * 4);
  return CFD_OK;
}

uint64_t cfd_mesh_hash(const cfd_mesh* m) {
  if (!m) return 0;
  uint64_t h = 0xcbf29ce484222325ull;
  auto feed = [&](const void* p, size_t n) {
    const unsigned char* b = (const unsigned char*)p;
    for (size_t i = 0; i < n; ++i) { h ^= b[i]; h *= 0x100000001b3ull; }
  };
  feed(m->x.data(), m->x.size() * 8);
  feed(m->y.data(), m->y.size() * 8);
  feed(m->cell_ptr.data(), m->cell_ptr.size() * 4);
  feed(m->cell_nodes.data(), m->cell_nodes.size() * 4);
  feed(m->bedges.data(), m->bedges.size() * 4);
  return h;
}

int cfd_mesh_get_lines(const cfd_mesh* m, int* n_segments, int* n_colors, int* seg_ptr, int* seg_cells,
                       int* segcol_ptr) {
  if (!m || !n_segments || !n_colors) return CFD_ERR_ARG;
  const int ns = (int)m->seg_ptr.size() - 1, nc = (int)m->segcol_ptr.size() - 1;
  *n_segments = ns;
  *n_colors = nc;
  if (seg_ptr) std::copy(m->seg_ptr.begin(), m->seg_ptr.end(), seg_ptr);
  if (seg_cells)
    for (size_t i = 0; i < m->seg_cells.size(); ++i) seg_cells[i] = m->iperm[m->seg_cells[i]];
  if (segcol_ptr) std::copy(m->segcol_ptr.begin(), m->segcol_ptr.end(), segcol_ptr);
  return CFD_OK;
}

int cfd_mesh_get_cell_colors(const cfd_mesh* m, int* n_colors, int* color) {
  if (!m || !n_colors) return CFD_ERR_ARG;
  *n_colors = (int)m->ccell_ptr.size() - 1;
  if (color)
    for (size_t i = 0; i < m->cell_color.size(); ++i) color[m->iperm[i]] = m->cell_color[i];
  return CFD_OK;
}

int cfd_mesh_write_gmsh(const cfd_mesh* m, const char* path) {
  if (!m || !path) return CFD_ERR_ARG;
  FILE* f = std::fopen(path, "w");
  if (!f) return CFD_ERR_IO;
  std::fprintf(f, "$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n%d\n", m->n_nodes);
  for (int i = 1; i <= m->n_nodes; ++i) std::fprintf(f, "%d %.17g %.17g 0\n", i, m->x[i], m->y[i]);
  int nb = (int)m->bedges.size() / 3, nc = m->ncells();
  std::fprintf(f, "$EndNodes\n$Elements\n%d\n", nb + nc);
  int id = 1;
  for (int e = 0; e < nb; ++e)
    std::fprintf(f, "%d 1 2 %d %d %d %d\n", id++, m->bedges[3 * e + 2], m->bedges[3 * e + 2], m->bedges[3 * e],
                 m->bedges[3 * e + 1]);
  for (int c = 0; c < nc; ++c) {
    int b = m->cell_ptr[c], n = m->cell_ptr[c + 1] - b;
    std::fprintf(f, "%d %d 2 100 100", id++, n == 3 ? 2 : 3);
    for (int k = 0; k < n; ++k) std::fprintf(f, " %d", m->cell_nodes[b + k]);
    std::fprintf(f, "\n");
  }
  std::fprintf(f, "$EndElements\n");
  std::fclose(f);
  return CFD_OK;
}

// main.cpp saveResults: the stream's default formatting of a double is "%g"
// at 6 significant digits, reproduced with printf.
int cfd_mesh_write_vtk(const cfd_mesh* m, const double* U6, double gamma, const char* path) {
  if (!m || !U6 || !path) return CFD_ERR_ARG;
  FILE* f = std::fopen(path, "w");
  if (!f) return CFD_ERR_IO;
  const int nc = m->ncells();
  std::fprintf(f, "# vtk DataFile Version 3.0\n2D Bump\nASCII\nDATASET UNSTRUCTURED_GRID\n");
  std::fprintf(f, "POINTS %d double\n", m->n_nodes);
  for (int i = 1; i <= m->n_nodes; ++i) std::fprintf(f, "%.6g %.6g 0.0\n", m->x[i], m->y[i]);
  long list = 0;
  for (int c = 0; c < nc; ++c) list += m->cell_ptr[c + 1] - m->cell_ptr[c] + 1;
  std::fprintf(f, "CELLS %d %ld\n", nc, list);
  for (int c = 0; c < nc; ++c) {
    const int b = m->cell_ptr[c], n = m->cell_ptr[c + 1] - b;
    std::fprintf(f, "%d", n);
    for (int k = 0; k < n; ++k) std::fprintf(f, " %d", m->cell_nodes[b + k] - 1);
    std::fprintf(f, "\n");
  }
  std::fprintf(f, "CELL_TYPES %d\n", nc);
  for (int c = 0; c < nc; ++c) std::fprintf(f, m->cell_ptr[c + 1] - m->cell_ptr[c] == 3 ? "5\n" : "9\n");
  std::fprintf(f, "CELL_DATA %d\n", nc);
  auto pres = [&](const double* u) {  // Flux.hpp getPressure
    const double V2 = (u[1] * u[1] + u[2] * u[2]) / (u[0] * u[0]);
    return (gamma - 1.0) * (u[3] - 0.5 * u[0] * V2);
  };
  std::fprintf(f, "SCALARS Mach double 1\nLOOKUP_TABLE default\n");
  for (int c = 0; c < nc; ++c) {
    const double* u = U6 + (size_t)6 * c;
    const double p = pres(u);
    const double a = std::sqrt(gamma * p / u[0]);
    const double v = std::sqrt(u[1] * u[1] + u[2] * u[2]) / u[0];
    std::fprintf(f, "%.6g\n", v / a);
  }
  std::fprintf(f, "SCALARS Pressure double 1\nLOOKUP_TABLE default\n");
  for (int c = 0; c < nc; ++c) std::fprintf(f, "%.6g\n", pres(U6 + (size_t)6 * c));
  std::fprintf(f, "SCALARS Density double 1\nLOOKUP_TABLE default\n");
  for (int c = 0; c < nc; ++c) std::fprintf(f, "%.6g\n", U6[(size_t)6 * c]);
  const bool ok = std::ferror(f) == 0;
  std::fclose(f);
  return ok ? CFD_OK : CFD_ERR_IO;
}

// Seeded synthetic meshes.  Nodes on an (nx+1) x (ny+1) grid, id = 1 + j (nx+1) + i.
// Quads visited row-major (j outer, i inner); stream 1 draws u per quad, the quad
// is split when u < p and a second draw's top bit picks the diagonal; a split
// quad is emitted as two consecutive counter-clockwise triangles.
int cfd_mesh_generate(int config, double scale, cfd_mesh** out) {
  if (!out || config < 1 || config > 5 || !(scale > 0)) return CFD_ERR_ARG;
  static const GenSpec specs[6] = {{0, 0, 0}, {160, 80, 1.0}, {400, 160, 0.2}, {1200, 1000, 0.1},
                                   {1280, 560, 0.15}, {2000, 2000, 0.1}};
  GenSpec s = specs[config];
  int nx = std::max(2, (int)std::lround(s.nx * scale));
  int ny = std::max(2, (int)std::lround(s.ny * scale));
  std::unique_ptr<cfd_mesh> m(new cfd_mesh());
  m->config = config;
  m->n_nodes = (nx + 1) * (ny + 1);
  m->x.assign(m->n_nodes + 1, 0.0);
  m->y.assign(m->n_nodes + 1, 0.0);
  auto nid = [&](int i, int j) { return 1 + j * (nx + 1) + i; };
  const double PI = 3.14159265358979323846;
  std::vector<double> xs(nx + 1), eta(ny + 1);
  // per-config node coordinates and boundary tags
  // tags: bottom(i) / top(i) / left / right
  std::vector<int> tb(nx), tt(nx);
  int tl = 1, tr = 2;
  if (config == 1) {  // 15-degree wedge ramp in [0,1] x [0,0.5]
    for (int i = 0; i <= nx; ++i) xs[i] = (double)i / nx;
    for (int j = 0; j <= ny; ++j) eta[j] = (double)j / ny;
    double tn = std::tan(15.0 * PI / 180.0);
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i <= nx; ++i) {
        double x = xs[i], yb = std::max(0.0, (x - 0.25) * tn);
        m->x[nid(i, j)] = x;
        m->y[nid(i, j)] = yb + (0.5 - yb) * eta[j];
      }
    for (int i = 0; i < nx; ++i) { tb[i] = 3; tt[i] = 1; }
  } else if (config == 2 || config == 3 || config == 5) {  // flat plate, x in [-0.1, 1]
    double H = 0.5, d1, r;
    if (config == 2) { d1 = 1.0 / 3240.0; r = 1.08; }
    else if (config == 3) { d1 = 4e-6; r = 1.1; }
    else { d1 = 2e-5; r = 1.05; }
    std::vector<double> yn = stretched(ny, H, d1, r);
    for (int i = 0; i <= nx; ++i) xs[i] = -0.1 + 1.1 * i / nx;
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i <= nx; ++i) {
        m->x[nid(i, j)] = xs[i];
        m->y[nid(i, j)] = yn[j];
      }
    for (int i = 0; i < nx; ++i) {
      double xmid = 0.5 * (xs[i] + xs[i + 1]);
      tb[i] = xmid < 0 ? 3 : 4;
      tt[i] = 5;
    }
    if (config == 5) {
      // oblique shock of 8 degrees deflection, Mach 2, hits the wall at x = 0.5:
      // top faces downstream of the shock foot get the post-shock state (tag 6)
      double beta = 0;
      {
        double M = 2.0, th = 8.0 * PI / 180.0, g = 1.4, lo = std::asin(1.0 / M) + 1e-9, hi = PI / 2;
        // weak-shock branch of tan(th) = 2 cot(b) (M^2 sin^2 b - 1) / (M^2 (g + cos 2b) + 2)
        auto fn = [&](double b) {
          return 2.0 / std::tan(b) * (M * M * std::sin(b) * std::sin(b) - 1.0) /
                     (M * M * (g + std::cos(2 * b)) + 2.0) - std::tan(th);
        };
        double bmax = lo, fmax = -1e30;
        for (int k = 0; k <= 2000; ++k) {
          double b = lo + (hi - lo) * k / 2000.0;
          if (fn(b) > fmax) { fmax = fn(b); bmax = b; }
        }
        double a = lo, b = bmax;
        for (int it = 0; it < 200; ++it) {
          double mid = 0.5 * (a + b);
          if (fn(mid) < 0) a = mid; else b = mid;
        }
        beta = 0.5 * (a + b);
      }
      double xtop = 0.5 - H / std::tan(beta);
      for (int i = 0; i < nx; ++i) {
        double xmid = 0.5 * (xs[i] + xs[i + 1]);
        tt[i] = xmid >= xtop ? 6 : 5;
      }
    }
  } else {  // config 4: 3 x 1 channel with a 5%-height circular-arc bump on [1,2]
    double h = 0.05, R = (0.25 + h * h) / (2 * h);
    for (int i = 0; i <= nx; ++i) xs[i] = 3.0 * i / nx;
    for (int j = 0; j <= ny; ++j) eta[j] = (double)j / ny;
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i <= nx; ++i) {
        double x = xs[i], yb = 0.0;
        if (x > 1.0 && x < 2.0) yb = std::sqrt(R * R - (x - 1.5) * (x - 1.5)) - (R - h);
        m->x[nid(i, j)] = x;
        m->y[nid(i, j)] = yb + (1.0 - yb) * eta[j];
      }
    for (int i = 0; i < nx; ++i) { tb[i] = 3; tt[i] = 3; }
  }
  SplitMix64 rng(stream_seed(config, 1));
  m->cell_nodes.reserve((size_t)nx * ny * 6);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      int a = nid(i, j), b = nid(i + 1, j), c = nid(i + 1, j + 1), d = nid(i, j + 1);
      double u = rng.uniform();
      if (u < s.split_p) {
        bool diag = (rng.next() >> 63) != 0;
        int t1[3], t2[3];
        if (!diag) { t1[0] = a; t1[1] = b; t1[2] = c; t2[0] = a; t2[1] = c; t2[2] = d; }
        else { t1[0] = a; t1[1] = b; t1[2] = d; t2[0] = b; t2[1] = c; t2[2] = d; }
        m->cell_nodes.insert(m->cell_nodes.end(), t1, t1 + 3);
        m->cell_ptr.push_back((int)m->cell_nodes.size());
        m->cell_nodes.insert(m->cell_nodes.end(), t2, t2 + 3);
        m->cell_ptr.push_back((int)m->cell_nodes.size());
      } else {
        int q[4] = {a, b, c, d};
        m->cell_nodes.insert(m->cell_nodes.end(), q, q + 4);
        m->cell_ptr.push_back((int)m->cell_nodes.size());
      }
    }
  auto addb = [&](int p, int q, int t) { m->bedges.push_back(p); m->bedges.push_back(q); m->bedges.push_back(t); };
  for (int i = 0; i < nx; ++i) addb(nid(i, 0), nid(i + 1, 0), tb[i]);
  for (int j = 0; j < ny; ++j) addb(nid(nx, j), nid(nx, j + 1), tr);
  for (int i = 0; i < nx; ++i) addb(nid(i, ny), nid(i + 1, ny), tt[i]);
  for (int j = 0; j < ny; ++j) addb(nid(0, j), nid(0, j + 1), tl);
  std::string err;
  int rc = cfdmesh::finish(*m, err);
  if (rc) return rc;
  *out = m.release();
  return CFD_OK;
}

}  // extern "C"
