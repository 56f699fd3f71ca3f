// Host mesh / DoF / constraint layer (deal.II-free) for axis-aligned hex
// meshes of FE_Q(1) or FE_Q(2) x 18 components.  It produces the data the
// reference builds in make_grid()/setup_system() (femgl/src/makegrid_*.cc,
// femgl/src/setup_uniform_B-phase.cc:125-198) and uploads once per cycle:
// cell->node maps, geometry, boundary ids, Dirichlet component masks,
// colour sets for the conflict-free scatter, and the BSR(18x18) pattern.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "common.hpp"

namespace glh {

struct Mesh {
  int deg = 1, n1 = 2, ns = 8;
  int n_cells = 0, n_nodes = 0;
  std::vector<int> cell_nodes;      // [n_cells*ns], lexicographic local order
  std::vector<double> cell_h;       // [n_cells*3]
  std::vector<double> cell_lo;      // [n_cells*3] lower corner
  std::vector<double> node_xyz;     // [n_nodes*3]
  std::vector<int> faces;           // boundary faces: (cell, local face 0..5, boundary id)
  std::vector<uint32_t> node_mask;  // [n_nodes] bit c = component c Dirichlet-constrained
  std::vector<int> color_ptr, color_cells;
  std::vector<int> row_ptr, col;    // BSR pattern over nodes, sorted columns
  // instances of a batch: instance i owns nodes [node_off[i], node_off[i+1])
  // and cells [cell_off[i], cell_off[i+1]); a single mesh is one instance
  std::vector<int> node_off, cell_off;
  // Adaptive meshes.  Hanging node h: U[h] = sum_k hang_w[k] U[hang_m[k]] over
  // k in [hang_ptr[h], hang_ptr[h+1]) (closed lines: no master is hanging).
  // row_ptr/col is then the condensed pattern of the system matrix
  // (constrained couplings moved to the masters, hanging rows diagonal only)
  // and arow_ptr/acol the pattern of the cell loop over all nodes.  All four
  // hang/arow vectors are empty for conforming meshes.
  std::vector<int> hang_ptr, hang_m;
  std::vector<double> hang_w;
  std::vector<int> arow_ptr, acol;
  std::vector<int> cell_level;  // refinement level of each cell (adaptive meshes)
  bool has_hanging() const { return !hang_ptr.empty() && hang_ptr.back() > 0; }
};

// Dirichlet component mask of a boundary id: 0 -> all 18 components (A = 0
// wall), 2/3/4 -> the x/y/z-normal orbital column (femgl/inc/femgl.h:294-301),
// anything else -> none.
uint32_t dirichlet_mask(int bid);

// Structured box mesh.  side_id for faces -x,+x,-y,+y,-z,+z.
int make_box(Mesh &m, int deg, int nx, int ny, int nz, const double lo[3], const double hi[3],
             const int side_id[6], std::string &err);

// BSR pattern: node pairs sharing a cell (DoFTools::make_sparsity_pattern
// with full component coupling, setup_uniform_B-phase.cc:190-196).
void build_pattern(Mesh &m);
// With hanging nodes: the cell-loop pattern into arow_ptr/acol and the
// condensed pattern (make_sparsity_pattern with constraints,
// keep_constrained_dofs = false) into row_ptr/col.
void build_constrained_patterns(Mesh &m);

// Greedy colouring of cells so that no two cells of one colour share a node.
void greedy_colors(Mesh &m);

// Mesh from a caller's arrays (the per-cycle data a deal.II caller has after
// make_grid()/setup_system(), setup_uniform_B-phase.cc:125-198): cells in
// the caller's node numbering, local order lexicographic (order 0) or FE_Q's
// hierarchic support-point order (order 1: vertices, lines, faces, interior
// in GeometryInfo<3> numbering), box geometry, boundary faces, optional
// Dirichlet component masks and node constraint lines (row node, masters,
// weights; closed here).  Colours and both BSR patterns are derived.
struct ArrayMesh {
  int degree = 1, n_cells = 0, n_nodes = 0, order = 0;
  const int *cell_nodes = nullptr;
  const double *cell_lo = nullptr, *cell_h = nullptr, *node_xyz = nullptr;
  int n_faces = 0;
  const int *faces = nullptr;
  const uint32_t *node_mask = nullptr;
  int n_lines = 0;
  const int *line_node = nullptr, *line_ptr = nullptr, *line_master = nullptr;
  const double *line_w = nullptr;
};
int from_arrays(Mesh &m, const ArrayMesh &a, std::string &err);
// local index (lexicographic, x fastest) of FE_Q(deg) support point i in
// hierarchic order
int hierarchic_to_lex(int deg, int i);

// Concatenate meshes of one degree into a batch with disjoint node and cell
// ranges; colour k of the batch is the union of colour k of every instance.
int concat(Mesh &out, const std::vector<const Mesh *> &ms, std::string &err);

// Index of block (r, s) in the pattern, or -1.
int find_block(const Mesh &m, int r, int s);
// Same in the cell-loop pattern (arow_ptr/acol when present).
int find_ablock(const Mesh &m, int r, int s);

}  // namespace glh
