// Adaptive hex meshes without deal.II: a forest of octrees over a box of
// root cells, refinement of flagged leaves with 2:1 balance, and the Q1 mesh
// of the leaves with closed hanging-node constraint lines.
//
// It stands in for the parallel::distributed::Triangulation work of the
// reference's adaptive cycle (femgl/src/refine.cc:133-173:
// refine_and_coarsen_fixed_number, execute_coarsening_and_refinement,
// SolutionTransfer::interpolate) and for
// DoFTools::make_hanging_node_constraints in setup_system()
// (femgl/src/setup_uniform_B-phase.cc:137-160).
//
// Leaves are kept in Morton order, so that cells and nodes that are close in
// space are close in memory.  Coordinates are integers on the lattice of the
// finest level LMAX: a leaf of level l spans 2^(LMAX-l) lattice units.
#pragma once
#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

#include "mesh.hpp"

namespace glh {

struct Forest {
  static constexpr int LMAX = 12;
  int nx = 1, ny = 1, nz = 1;  // root cells
  double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
  int side_id[6] = {1, 1, 1, 1, 1, 1};
  struct Leaf {
    int level, ix, iy, iz;  // index among the cells of its level
  };
  std::vector<Leaf> leaves;
  std::unordered_map<uint64_t, int> where;  // key(level, index) -> leaf

  static uint64_t key(int l, int i, int j, int k)
  {
    return ((uint64_t)l << 57) | ((uint64_t)i << 38) | ((uint64_t)j << 19) | (uint64_t)k;
  }
  // leaf covering cell (i,j,k) of level l (at that level or coarser), or -1
  int locate(int l, int i, int j, int k) const;
  // axis d is periodic: side id 5 on both of its faces (the matched faces of
  // makegrid_retangle-z-AdGR_xy-periodic.cc); nodes on its + face are the
  // nodes of its - face, and neighbours wrap around
  bool periodic(int d) const { return side_id[2 * d] == 5 && side_id[2 * d + 1] == 5; }
  void reindex();  // Morton sort of the leaves and rebuild of `where`
};

int forest_init(Forest &f, int nx, int ny, int nz, const double lo[3], const double hi[3], const int side_id[6],
                std::string &err);
// refine_and_coarsen: flag 2 marks a leaf for coarsening, 1 for refinement.
// Complete families of 8 leaves all flagged 2 become their parent where the
// 2:1 balance allows it; then the leaves flagged 1 are refined, and refined
// further until neighbouring leaves (across faces, edges and vertices)
// differ by at most one level.
int forest_refine(Forest &f, const std::vector<char> &flag, std::string &err);
// Q1 or Q2 (deg) mesh of the leaves (cell e = leaf e).  Hanging nodes get closed
// constraint lines in m.hang_*, the flag HANG_BIT in node_mask and no
// Dirichlet bits; m.row_ptr/col is the condensed system pattern and
// m.arow_ptr/acol the pattern the cell loop assembles into.
int forest_mesh(const Forest &f, Mesh &m, std::string &err, int deg = 1);
// Lagrange basis function i of degree deg (1 or 2) on the equidistant nodes
// of [0,1], evaluated at t
double lagrange(int deg, int i, double t);
// The interpolation of forest_transfer: for each node of mn, the fo cell
// that contains it and the node's coordinates t[3] in [0,1]^3 within it
// (dyadic fractions of level <= LMAX, exact in float).
int forest_transfer_map(const Forest &fo, const Mesh &mo, const Forest &fn, const Mesh &mn, std::vector<int> &cell,
                        std::vector<float> &t, std::string &err);
// Interpolate a Q1 solution of the mesh of `fo` (mesh mo) to the mesh of
// `fn` (mesh mn), which refines it: every vertex of mn takes the trilinear
// interpolant of the leaf of fo that contains it (SolutionTransfer of Q1).
int forest_transfer(const Forest &fo, const Mesh &mo, const double *Uo, const Forest &fn, const Mesh &mn,
                    double *Un, std::string &err);

// Kelly indicators (one per leaf) of the Q1 field U on the mesh of f.
int kelly_estimate(const Forest &f, const Mesh &m, const double *U, float *eta, std::string &err);
// Flags for the floor(frac * n) cells of largest indicator (ties in cell order).
void flag_fixed_number(int n, const float *eta, double frac, std::vector<char> &flag);
// refine_and_coarsen_fixed_number: 1 for the floor(refine * n) cells of
// largest indicator, 2 for the floor(coarsen * n) of smallest (ties in cell
// order), 0 otherwise
void flag_refine_coarsen(int n, const float *eta, double refine, double coarsen, std::vector<char> &flag);

}  // namespace glh
