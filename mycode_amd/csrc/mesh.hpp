// Host-side mesh representation shared by the mesh pipeline and the device
// context.  Reference-order arrays reproduce Mesh::faces / Mesh::cells
// (Types.hpp:16-35) field for field; the reorder arrays describe the device
// layout (RCM cell order, faces grouped by edge color).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/cfd_gpu.h"

struct cfd_mesh {
  // description (Gmsh-like, 1-based node ids)
  int n_nodes = 0;
  std::vector<double> x, y;  // n_nodes + 1
  std::vector<int> cell_ptr{0}, cell_nodes;
  std::vector<int> bedges;  // n1, n2, tag triples
  // faces in reference order
  std::vector<int> n1, n2, own, nei, tag;
  std::vector<double> nx, ny, area, xm, ym;
  // cells in reference order
  std::vector<double> vol, xc, yc;
  // device layout
  std::vector<int> perm;        // reference cell -> RCM position
  std::vector<int> iperm;       // RCM position -> reference cell
  std::vector<int> color;       // reference face -> color
  int ncolors = 0;
  std::vector<int> face_order;  // device face slot -> reference face
  std::vector<int> color_ptr;   // ncolors + 1 offsets into face_order
  // distance-1 cell coloring (no two cells sharing a face alike), RCM positions
  std::vector<int> cell_color;  // per RCM position
  std::vector<int> ccells;      // RCM positions grouped by cell color, ascending within a color
  std::vector<int> ccell_ptr;   // n_cell_colors + 1
  // line segments of the line-implicit preconditioner (RCM positions): cells
  // grouped into chains along their strongest couplings, cut to at most 64
  // cells; segments are listed color by color (no two segments of one color
  // share a face), each segment's cells in chain order
  std::vector<int> seg_ptr;     // n_segments + 1 offsets into seg_cells
  std::vector<int> seg_cells;   // N cells
  std::vector<int> segcol_ptr;  // n_segment_colors + 1 offsets into the segment list
  std::vector<int> lslot_prev;  // per RCM position: slot of the previous cell of its segment, or -1
  std::vector<int> lslot_next;  // per RCM position: slot of the next cell of its segment, or -1
  // synthetic-mesh extras (valid after cfd_mesh_generate)
  int config = 0;
  std::vector<double> init_U6;  // seeded initial state, reference order (may be empty)

  int ncells() const { return (int)cell_ptr.size() - 1; }
  int nfaces() const { return (int)own.size(); }
};

namespace cfdmesh {
// Mesh::buildConnectivity + Mesh::computeGeometry + RCM + coloring
int finish(cfd_mesh& m, std::string& err);
}
