#include "mesh.hpp"

#include <algorithm>
#include <cmath>
#include <map>

namespace glh {

uint32_t dirichlet_mask(int bid)
{
  // 0: A = 0 on all components; 2/3/4: the x/y/z-normal orbital column;
  // 1 (natural) and 5 (periodic) constrain nothing here
  if (bid == 0) return (1u << NC) - 1u;
  if (bid >= 2 && bid <= 4) {
    uint32_t m = 0;
    int col = bid - 2;
    for (int c = 0; c < NC; ++c)
      if ((c % 9) % 3 == col) m |= 1u << c;
    return m;
  }
  return 0u;
}

int make_box(Mesh &m, int deg, int nx, int ny, int nz, const double lo[3], const double hi[3],
             const int side_id[6], std::string &err)
{
  if (deg != 1 && deg != 2) { err = "degree must be 1 or 2"; return -1; }
  if (nx < 1 || ny < 1 || nz < 1) { err = "cell counts must be positive"; return -1; }
  for (int k = 0; k < 3; ++k)
    if (!(hi[k] > lo[k])) { err = "empty box"; return -1; }
  m = Mesh();
  m.deg = deg; m.n1 = deg + 1; m.ns = m.n1 * m.n1 * m.n1;
  const long Nx = (long)deg * nx + 1, Ny = (long)deg * ny + 1, Nz = (long)deg * nz + 1;
  const long nn = Nx * Ny * Nz, nc = (long)nx * ny * nz;
  if (nn >= (1L << 31) / NC || nc * m.ns >= (1L << 31)) { err = "mesh too large for 32-bit indices"; return -1; }
  m.n_nodes = (int)nn; m.n_cells = (int)nc;
  const double h[3] = {(hi[0] - lo[0]) / nx, (hi[1] - lo[1]) / ny, (hi[2] - lo[2]) / nz};
  m.node_xyz.resize(3 * nn);
  for (long k = 0; k < Nz; ++k)
    for (long j = 0; j < Ny; ++j)
      for (long i = 0; i < Nx; ++i) {
        long n = i + Nx * (j + Ny * k);
        m.node_xyz[3 * n + 0] = lo[0] + (double)i * h[0] / deg;
        m.node_xyz[3 * n + 1] = lo[1] + (double)j * h[1] / deg;
        m.node_xyz[3 * n + 2] = lo[2] + (double)k * h[2] / deg;
      }
  m.cell_nodes.resize(nc * m.ns);
  m.cell_h.resize(3 * nc);
  m.cell_lo.resize(3 * nc);
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        long e = i + (long)nx * (j + (long)ny * k);
        int a = 0;
        for (int c = 0; c < m.n1; ++c)
          for (int b = 0; b < m.n1; ++b)
            for (int q = 0; q < m.n1; ++q, ++a)
              m.cell_nodes[e * m.ns + a] =
                  (int)(((long)deg * i + q) + Nx * (((long)deg * j + b) + Ny * ((long)deg * k + c)));
        for (int d = 0; d < 3; ++d) m.cell_h[3 * e + d] = h[d];
        m.cell_lo[3 * e + 0] = lo[0] + i * h[0];
        m.cell_lo[3 * e + 1] = lo[1] + j * h[1];
        m.cell_lo[3 * e + 2] = lo[2] + k * h[2];
        const bool on[6] = {i == 0, i == nx - 1, j == 0, j == ny - 1, k == 0, k == nz - 1};
        for (int f = 0; f < 6; ++f)
          if (on[f] && side_id[f] != 1) {
            m.faces.push_back((int)e);
            m.faces.push_back(f);
            m.faces.push_back(side_id[f]);
          }
      }
  // Dirichlet masks: nodes on a side take that side's component mask
  m.node_mask.assign(nn, 0u);
  for (long k = 0; k < Nz; ++k)
    for (long j = 0; j < Ny; ++j)
      for (long i = 0; i < Nx; ++i) {
        long n = i + Nx * (j + Ny * k);
        const bool on[6] = {i == 0, i == Nx - 1, j == 0, j == Ny - 1, k == 0, k == Nz - 1};
        uint32_t mk = 0;
        for (int f = 0; f < 6; ++f)
          if (on[f]) mk |= dirichlet_mask(side_id[f]);
        m.node_mask[n] = mk;
      }
  // colours: parity of the cell index (no two cells of one parity class share a node)
  std::vector<std::vector<int>> cls(8);
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) cls[(i & 1) + 2 * (j & 1) + 4 * (k & 1)].push_back(i + nx * (j + ny * k));
  m.color_ptr.assign(1, 0);
  for (auto &c : cls) {
    if (c.empty()) continue;
    m.color_cells.insert(m.color_cells.end(), c.begin(), c.end());
    m.color_ptr.push_back((int)m.color_cells.size());
  }
  // periodic axes (side id 5 on both faces, the matched face pairs of
  // makegrid_*-periodic.cc / make_periodicity_constraints): a node on a +
  // face takes the value of its image on the - faces (weight 1; corners and
  // edges map straight to the image on every periodic axis, so the lines are
  // closed)
  bool per[3];
  for (int k = 0; k < 3; ++k) {
    per[k] = side_id[2 * k] == 5 && side_id[2 * k + 1] == 5;
    if ((side_id[2 * k] == 5) != (side_id[2 * k + 1] == 5)) {
      err = "a periodic axis needs side id 5 on both faces";
      return -1;
    }
  }
  if (per[0] || per[1] || per[2]) {
    m.hang_ptr.assign(nn + 1, 0);
    for (long k = 0; k < Nz; ++k)
      for (long j = 0; j < Ny; ++j)
        for (long i = 0; i < Nx; ++i) {
          const long n = i + Nx * (j + Ny * k);
          const long mi = (per[0] && i == Nx - 1) ? 0 : i, mj = (per[1] && j == Ny - 1) ? 0 : j,
                     mk = (per[2] && k == Nz - 1) ? 0 : k;
          const long mn = mi + Nx * (mj + Ny * mk);
          if (mn != n) {
            m.hang_m.push_back((int)mn);
            m.hang_w.push_back(1.0);
            m.node_mask[n] = HANG_BIT;
          }
          m.hang_ptr[n + 1] = (int)m.hang_m.size();
        }
    build_constrained_patterns(m);
  } else {
    build_pattern(m);
  }
  m.node_off = {0, m.n_nodes};
  m.cell_off = {0, m.n_cells};
  return 0;
}

int concat(Mesh &out, const std::vector<const Mesh *> &ms, std::string &err)
{
  if (ms.empty()) { err = "empty batch"; return -1; }
  out = Mesh();
  out.deg = ms[0]->deg; out.n1 = ms[0]->n1; out.ns = ms[0]->ns;
  long nn = 0, nc = 0, nb = 0;
  size_t ncol = 0;
  for (const Mesh *m : ms) {
    if (m->deg != out.deg) { err = "all instances of a batch need the same degree"; return -1; }
    nn += m->n_nodes; nc += m->n_cells; nb += m->row_ptr.back();
    ncol = std::max(ncol, m->color_ptr.size() - 1);
  }
  if (nn >= (1L << 31) / NC || nc * out.ns >= (1L << 31) || nb >= (1L << 31)) {
    err = "batch too large for 32-bit indices";
    return -1;
  }
  out.n_nodes = (int)nn; out.n_cells = (int)nc;
  out.node_off.assign(1, 0);
  out.cell_off.assign(1, 0);
  out.row_ptr.assign(1, 0);
  int noff = 0, coff = 0;
  for (const Mesh *m : ms) {
    for (int v : m->cell_nodes) out.cell_nodes.push_back(v + noff);
    out.cell_h.insert(out.cell_h.end(), m->cell_h.begin(), m->cell_h.end());
    out.cell_lo.insert(out.cell_lo.end(), m->cell_lo.begin(), m->cell_lo.end());
    out.node_xyz.insert(out.node_xyz.end(), m->node_xyz.begin(), m->node_xyz.end());
    for (size_t f = 0; f + 2 < m->faces.size(); f += 3) {
      out.faces.push_back(m->faces[f] + coff);
      out.faces.push_back(m->faces[f + 1]);
      out.faces.push_back(m->faces[f + 2]);
    }
    out.node_mask.insert(out.node_mask.end(), m->node_mask.begin(), m->node_mask.end());
    const int base = out.row_ptr.back();
    for (int r = 0; r < m->n_nodes; ++r) out.row_ptr.push_back(base + m->row_ptr[r + 1]);
    for (int v : m->col) out.col.push_back(v + noff);
    noff += m->n_nodes;
    coff += m->n_cells;
    out.node_off.push_back(noff);
    out.cell_off.push_back(coff);
  }
  // hanging-node lines and cell-loop patterns, when any instance has them
  bool any_hang = false;
  for (const Mesh *m : ms) any_hang = any_hang || m->has_hanging();
  if (any_hang) {
    out.hang_ptr.assign(1, 0);
    out.arow_ptr.assign(1, 0);
    int nof = 0;
    for (const Mesh *m : ms) {
      const bool h = m->has_hanging();
      for (int v = 0; v < m->n_nodes; ++v) {
        if (h)
          for (int k = m->hang_ptr[v]; k < m->hang_ptr[v + 1]; ++k) {
            out.hang_m.push_back(m->hang_m[k] + nof);
            out.hang_w.push_back(m->hang_w[k]);
          }
        out.hang_ptr.push_back((int)out.hang_m.size());
      }
      const std::vector<int> &rp = h ? m->arow_ptr : m->row_ptr;
      const std::vector<int> &cl = h ? m->acol : m->col;
      const int base = out.arow_ptr.back();
      for (int r = 0; r < m->n_nodes; ++r) out.arow_ptr.push_back(base + rp[r + 1]);
      for (int v : cl) out.acol.push_back(v + nof);
      nof += m->n_nodes;
    }
  }
  out.color_ptr.assign(1, 0);
  for (size_t k = 0; k < ncol; ++k) {
    int cbase = 0;
    for (const Mesh *m : ms) {
      if (k + 1 < m->color_ptr.size())
        for (int i = m->color_ptr[k]; i < m->color_ptr[k + 1]; ++i) out.color_cells.push_back(m->color_cells[i] + cbase);
      cbase += m->n_cells;
    }
    out.color_ptr.push_back((int)out.color_cells.size());
  }
  return 0;
}

void build_pattern(Mesh &m)
{
  std::vector<std::vector<int>> nb(m.n_nodes);
  for (int e = 0; e < m.n_cells; ++e) {
    const int *cn = &m.cell_nodes[(size_t)e * m.ns];
    for (int a = 0; a < m.ns; ++a)
      for (int b = 0; b < m.ns; ++b) nb[cn[a]].push_back(cn[b]);
  }
  m.row_ptr.assign(m.n_nodes + 1, 0);
  m.col.clear();
  for (int r = 0; r < m.n_nodes; ++r) {
    auto &v = nb[r];
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    m.row_ptr[r + 1] = m.row_ptr[r] + (int)v.size();
    m.col.insert(m.col.end(), v.begin(), v.end());
    std::vector<int>().swap(v);
  }
}

void greedy_colors(Mesh &m)
{
  std::vector<std::vector<int>> node_cells(m.n_nodes);
  for (int e = 0; e < m.n_cells; ++e)
    for (int a = 0; a < m.ns; ++a) node_cells[m.cell_nodes[(size_t)e * m.ns + a]].push_back(e);
  std::vector<int> color(m.n_cells, -1);
  int ncol = 0;
  std::vector<int> used;
  for (int e = 0; e < m.n_cells; ++e) {
    used.assign(ncol + 1, 0);
    for (int a = 0; a < m.ns; ++a)
      for (int o : node_cells[m.cell_nodes[(size_t)e * m.ns + a]])
        if (color[o] >= 0) used[color[o]] = 1;
    int c = 0;
    while (c < ncol && used[c]) ++c;
    color[e] = c;
    if (c == ncol) ++ncol;
  }
  m.color_ptr.assign(ncol + 1, 0);
  for (int e = 0; e < m.n_cells; ++e) ++m.color_ptr[color[e] + 1];
  for (int c = 0; c < ncol; ++c) m.color_ptr[c + 1] += m.color_ptr[c];
  m.color_cells.assign(m.n_cells, 0);
  std::vector<int> pos(m.color_ptr.begin(), m.color_ptr.end() - 1);
  for (int e = 0; e < m.n_cells; ++e) m.color_cells[pos[color[e]]++] = e;
}

int find_block(const Mesh &m, int r, int s)
{
  auto b = m.col.begin() + m.row_ptr[r], e = m.col.begin() + m.row_ptr[r + 1];
  auto it = std::lower_bound(b, e, s);
  if (it == e || *it != s) return -1;
  return (int)(it - m.col.begin());
}

int find_ablock(const Mesh &m, int r, int s)
{
  if (m.arow_ptr.empty()) return find_block(m, r, s);
  auto b = m.acol.begin() + m.arow_ptr[r], e = m.acol.begin() + m.arow_ptr[r + 1];
  auto it = std::lower_bound(b, e, s);
  if (it == e || *it != s) return -1;
  return (int)(it - m.acol.begin());
}

void build_constrained_patterns(Mesh &m)
{
  build_pattern(m);
  m.arow_ptr.swap(m.row_ptr);
  m.acol.swap(m.col);
  // targets of a node: itself, or the masters of its constraint line
  auto targets = [&](int a, std::vector<int> &t) {
    t.clear();
    if (m.hang_ptr[a + 1] > m.hang_ptr[a])
      for (int k = m.hang_ptr[a]; k < m.hang_ptr[a + 1]; ++k) t.push_back(m.hang_m[k]);
    else
      t.push_back(a);
  };
  std::vector<std::vector<int>> nb(m.n_nodes);
  std::vector<int> ta, tb;
  for (int e = 0; e < m.n_cells; ++e) {
    const int *cn = &m.cell_nodes[(size_t)e * m.ns];
    for (int a = 0; a < m.ns; ++a) {
      targets(cn[a], ta);
      for (int b = 0; b < m.ns; ++b) {
        targets(cn[b], tb);
        for (int r : ta) nb[r].insert(nb[r].end(), tb.begin(), tb.end());
      }
    }
  }
  for (int r = 0; r < m.n_nodes; ++r) nb[r].push_back(r);  // every row keeps its diagonal block
  m.row_ptr.assign(m.n_nodes + 1, 0);
  m.col.clear();
  for (int r = 0; r < m.n_nodes; ++r) {
    auto &v = nb[r];
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    m.row_ptr[r + 1] = m.row_ptr[r] + (int)v.size();
    m.col.insert(m.col.end(), v.begin(), v.end());
    std::vector<int>().swap(v);
  }
}

}  // namespace glh

namespace glh {

int hierarchic_to_lex(int deg, int i)
{
  // FE_Q support points in the order deal.II numbers their dofs: the 8
  // vertices (x fastest), then for Q2 the midpoints of the 12 lines, of the
  // 6 faces and of the cell, with lines and faces numbered as GeometryInfo<3>
  // does (lines 0-3 on z = 0: x = 0, x = 1, y = 0, y = 1; lines 4-7 the same
  // on z = 1; lines 8-11 along z at (0,0), (1,0), (0,1), (1,1); faces -x, +x,
  // -y, +y, -z, +z).  Entries are (x, y, z) in units of half a cell.
  static const int q2[27][3] = {
      {0, 0, 0}, {2, 0, 0}, {0, 2, 0}, {2, 2, 0}, {0, 0, 2}, {2, 0, 2}, {0, 2, 2}, {2, 2, 2},  // vertices
      {0, 1, 0}, {2, 1, 0}, {1, 0, 0}, {1, 2, 0},                                              // lines 0-3
      {0, 1, 2}, {2, 1, 2}, {1, 0, 2}, {1, 2, 2},                                              // lines 4-7
      {0, 0, 1}, {2, 0, 1}, {0, 2, 1}, {2, 2, 1},                                              // lines 8-11
      {0, 1, 1}, {2, 1, 1}, {1, 0, 1}, {1, 2, 1}, {1, 1, 0}, {1, 1, 2},                        // faces
      {1, 1, 1}};                                                                              // interior
  if (deg == 1) return i;  // the vertices are already lexicographic
  return q2[i][0] + 3 * (q2[i][1] + 3 * q2[i][2]);
}

int from_arrays(Mesh &m, const ArrayMesh &a, std::string &err)
{
  if (a.degree != 1 && a.degree != 2) { err = "degree must be 1 or 2"; return -1; }
  if (a.order != 0 && a.order != 1) { err = "local order must be 0 (lexicographic) or 1 (hierarchic)"; return -1; }
  if (a.n_cells < 1 || a.n_nodes < 1) { err = "a mesh needs cells and nodes"; return -1; }
  if (!a.cell_nodes || !a.cell_lo || !a.cell_h) { err = "cell_nodes, cell_lo and cell_h are required"; return -1; }
  if (a.n_faces < 0 || (a.n_faces > 0 && !a.faces)) { err = "bad face list"; return -1; }
  if (a.n_lines < 0 || (a.n_lines > 0 && (!a.line_node || !a.line_ptr || !a.line_master || !a.line_w))) {
    err = "bad constraint line arrays";
    return -1;
  }
  const int deg = a.degree, n1 = deg + 1, ns = n1 * n1 * n1;
  const long nn = a.n_nodes, nc = a.n_cells;
  if (nn >= (1L << 31) / NC || nc * ns >= (1L << 31)) { err = "mesh too large for 32-bit indices"; return -1; }
  m = Mesh();
  m.deg = deg; m.n1 = n1; m.ns = ns;
  m.n_cells = (int)nc; m.n_nodes = (int)nn;
  m.cell_nodes.resize(nc * ns);
  m.cell_h.assign(a.cell_h, a.cell_h + 3 * nc);
  m.cell_lo.assign(a.cell_lo, a.cell_lo + 3 * nc);
  std::vector<char> used(nn, 0);
  for (long e = 0; e < nc; ++e) {
    for (int k = 0; k < 3; ++k)
      if (!(m.cell_h[3 * e + k] > 0.0) || !std::isfinite(m.cell_h[3 * e + k]) || !std::isfinite(m.cell_lo[3 * e + k])) {
        err = "cell " + std::to_string(e) + ": edge lengths must be positive and finite";
        return -1;
      }
    int *cn = &m.cell_nodes[e * ns];
    for (int i = 0; i < ns; ++i) {
      const int v = a.cell_nodes[e * ns + i];
      if (v < 0 || v >= nn) { err = "cell " + std::to_string(e) + ": node index out of range"; return -1; }
      cn[a.order == 1 ? hierarchic_to_lex(deg, i) : i] = v;
      used[v] = 1;
    }
    for (int i = 0; i < ns; ++i)
      for (int j = i + 1; j < ns; ++j)
        if (cn[i] == cn[j]) { err = "cell " + std::to_string(e) + ": repeated node"; return -1; }
  }
  for (long v = 0; v < nn; ++v)
    if (!used[v]) { err = "node " + std::to_string(v) + " belongs to no cell"; return -1; }
  // support points: every cell must put its nodes where the caller says the
  // nodes are (or, without coordinates, where the first cell puts them),
  // which also catches a wrong local order
  std::vector<char> seen(nn, 0);
  m.node_xyz.assign(3 * nn, 0.0);
  if (a.node_xyz) {
    for (long k = 0; k < 3 * nn; ++k)
      if (!std::isfinite(a.node_xyz[k])) { err = "non-finite node coordinate"; return -1; }
    m.node_xyz.assign(a.node_xyz, a.node_xyz + 3 * nn);
  }
  for (long e = 0; e < nc; ++e) {
    const double *lo = &m.cell_lo[3 * e], *h = &m.cell_h[3 * e];
    const double tol = 1e-9 * (std::fabs(lo[0]) + std::fabs(lo[1]) + std::fabs(lo[2]) + h[0] + h[1] + h[2]);
    for (int i = 0; i < ns; ++i) {
      const int v = m.cell_nodes[e * ns + i];
      const int t[3] = {i % n1, (i / n1) % n1, i / (n1 * n1)};
      double P[3];
      for (int k = 0; k < 3; ++k) P[k] = lo[k] + h[k] * t[k] / deg;
      if (!a.node_xyz && !seen[v]) {
        for (int k = 0; k < 3; ++k) m.node_xyz[3 * v + k] = P[k];
        seen[v] = 1;
        continue;
      }
      for (int k = 0; k < 3; ++k)
        if (std::fabs(m.node_xyz[3 * v + k] - P[k]) > tol) {
          err = "cell " + std::to_string(e) + ": node " + std::to_string(v) +
                " is not at the cell's support point (wrong local order or coordinates?)";
          return -1;
        }
    }
  }
  // boundary faces; natural faces (id 1) carry nothing
  for (int f = 0; f < a.n_faces; ++f) {
    const int e = a.faces[3 * f], lf = a.faces[3 * f + 1], bid = a.faces[3 * f + 2];
    if (e < 0 || e >= nc || lf < 0 || lf > 5 || bid < 0) { err = "face " + std::to_string(f) + " out of range"; return -1; }
    if (bid == 1) continue;
    m.faces.push_back(e);
    m.faces.push_back(lf);
    m.faces.push_back(bid);
  }
  // Dirichlet component masks: the caller's, or the union of the masks of
  // the boundary ids of the faces a node lies on (interpolate_boundary_values
  // with the component masks of femgl.h:294-301)
  m.node_mask.assign(nn, 0u);
  if (a.node_mask) {
    for (long v = 0; v < nn; ++v) {
      if (a.node_mask[v] >> NC) { err = "node mask bits above component 17"; return -1; }
      m.node_mask[v] = a.node_mask[v];
    }
  } else {
    for (size_t f = 0; f + 2 < m.faces.size(); f += 3) {
      const int e = m.faces[f], lf = m.faces[f + 1];
      const uint32_t mk = dirichlet_mask(m.faces[f + 2]);
      if (!mk) continue;
      const int ax = lf / 2, side = (lf % 2) * deg;
      for (int i = 0; i < ns; ++i) {
        const int t[3] = {i % n1, (i / n1) % n1, i / (n1 * n1)};
        if (t[ax] == side) m.node_mask[m.cell_nodes[(size_t)e * ns + i]] |= mk;
      }
    }
  }
  // constraint lines (hanging nodes, periodic images), closed by substituting
  // constrained masters (AffineConstraints::close)
  if (a.n_lines > 0) {
    std::vector<std::vector<std::pair<int, double>>> line(nn);
    std::vector<char> con(nn, 0);
    for (int l = 0; l < a.n_lines; ++l) {
      const int v = a.line_node[l], k0 = a.line_ptr[l], k1 = a.line_ptr[l + 1];
      if (v < 0 || v >= nn) { err = "constraint line " + std::to_string(l) + ": node out of range"; return -1; }
      if (con[v]) { err = "node " + std::to_string(v) + " has two constraint lines"; return -1; }
      if ((l == 0 && k0 != 0) || k1 <= k0) { err = "line_ptr must start at 0 and increase"; return -1; }
      con[v] = 1;
      for (int k = k0; k < k1; ++k) {
        const int u = a.line_master[k];
        const double w = a.line_w[k];
        if (u < 0 || u >= nn || u == v || !std::isfinite(w)) {
          err = "constraint line " + std::to_string(l) + ": bad master or weight";
          return -1;
        }
        line[v].push_back({u, w});
      }
    }
    for (long v = 0; v < nn; ++v) {
      if (!con[v]) continue;
      for (int pass = 0;; ++pass) {
        if (pass == 64) { err = "cyclic constraint lines"; return -1; }
        bool changed = false;
        std::map<int, double> acc;
        for (auto &mw : line[v]) {
          if (con[mw.first]) {
            if (mw.first == v) { err = "cyclic constraint lines"; return -1; }
            for (auto &mw2 : line[mw.first]) acc[mw2.first] += mw.second * mw2.second;
            changed = true;
          } else {
            acc[mw.first] += mw.second;
          }
        }
        line[v].assign(acc.begin(), acc.end());
        if (!changed) break;
      }
    }
    m.hang_ptr.assign(nn + 1, 0);
    for (long v = 0; v < nn; ++v) {
      m.hang_ptr[v + 1] = m.hang_ptr[v] + (con[v] ? (int)line[v].size() : 0);
      if (!con[v]) continue;
      m.node_mask[v] = HANG_BIT;  // the line wins over a Dirichlet mask, as it is added first
      for (auto &mw : line[v]) {
        m.hang_m.push_back(mw.first);
        m.hang_w.push_back(mw.second);
      }
    }
  }
  greedy_colors(m);
  if (m.has_hanging()) build_constrained_patterns(m);
  else {
    build_pattern(m);
    m.hang_ptr.clear();
  }
  m.node_off = {0, m.n_nodes};
  m.cell_off = {0, m.n_cells};
  return 0;
}

}  // namespace glh
