#include "amr.hpp"

#include <algorithm>
#include <cmath>
#include <map>

namespace glh {

namespace {

// Morton code of a lattice point (21 bits per coordinate)
uint64_t morton(uint64_t x, uint64_t y, uint64_t z)
{
  auto spread = [](uint64_t v) {
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
  };
  return spread(x) | spread(y) << 1 | spread(z) << 2;
}

uint64_t vkey(int64_t x, int64_t y, int64_t z) { return ((uint64_t)x << 42) | ((uint64_t)y << 21) | (uint64_t)z; }

}  // namespace

int Forest::locate(int l, int i, int j, int k) const
{
  for (int lv = l; lv >= 0; --lv) {
    const int s = l - lv;
    auto it = where.find(key(lv, i >> s, j >> s, k >> s));
    if (it != where.end()) return it->second;
  }
  return -1;
}

void Forest::reindex()
{
  std::vector<std::pair<uint64_t, int>> ord(leaves.size());
  for (size_t e = 0; e < leaves.size(); ++e) {
    const Leaf &L = leaves[e];
    const int s = LMAX - L.level;
    // lower corner on the finest lattice; ties (same corner) ordered by level
    ord[e] = {morton((uint64_t)L.ix << s, (uint64_t)L.iy << s, (uint64_t)L.iz << s) * 16 + (uint64_t)L.level, (int)e};
  }
  std::sort(ord.begin(), ord.end());
  std::vector<Leaf> nl(leaves.size());
  for (size_t e = 0; e < ord.size(); ++e) nl[e] = leaves[ord[e].second];
  leaves.swap(nl);
  where.clear();
  where.reserve(leaves.size() * 2);
  for (size_t e = 0; e < leaves.size(); ++e) where[key(leaves[e].level, leaves[e].ix, leaves[e].iy, leaves[e].iz)] = (int)e;
}

int forest_init(Forest &f, int nx, int ny, int nz, const double lo[3], const double hi[3], const int side_id[6],
                std::string &err)
{
  if (nx < 1 || ny < 1 || nz < 1) { err = "cell counts must be positive"; return -1; }
  if (((int64_t)std::max(nx, std::max(ny, nz)) << Forest::LMAX) >= (1 << 19)) {
    err = "too many root cells for the refinement lattice";
    return -1;
  }
  for (int k = 0; k < 3; ++k)
    if (!(hi[k] > lo[k])) { err = "empty box"; return -1; }
  const int nroot[3] = {nx, ny, nz};
  for (int k = 0; k < 3; ++k) {
    if ((side_id[2 * k] == 5) != (side_id[2 * k + 1] == 5)) {
      err = "a periodic axis needs side id 5 on both faces";
      return -1;
    }
    if (side_id[2 * k] == 5 && nroot[k] < 2) {
      err = "a periodic axis needs at least two root cells";
      return -1;
    }
  }
  f = Forest();
  f.nx = nx; f.ny = ny; f.nz = nz;
  for (int k = 0; k < 3; ++k) { f.lo[k] = lo[k]; f.hi[k] = hi[k]; }
  for (int k = 0; k < 6; ++k) f.side_id[k] = side_id[k];
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) f.leaves.push_back({0, i, j, k});
  f.reindex();
  return 0;
}

namespace {

// 2:1 balance violations: leaves that have a neighbour (26 directions, across
// periodic faces too) more than one level finer; returned as leaf indices
std::vector<int> too_coarse(const Forest &f)
{
  std::vector<int> out;
  std::vector<char> seen(f.leaves.size(), 0);
  for (const auto &L : f.leaves) {
    if (L.level < 2) continue;
    const int n[3] = {f.nx << L.level, f.ny << L.level, f.nz << L.level};
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          if (!dx && !dy && !dz) continue;
          int q[3] = {L.ix + dx, L.iy + dy, L.iz + dz};
          bool outside = false;
          for (int d = 0; d < 3; ++d) {
            if (f.periodic(d)) q[d] = (q[d] + n[d]) % n[d];  // across the periodic faces
            else if (q[d] < 0 || q[d] >= n[d]) outside = true;
          }
          if (outside) continue;
          const int K = f.locate(L.level, q[0], q[1], q[2]);
          if (K >= 0 && f.leaves[K].level < L.level - 1 && !seen[K]) {
            seen[K] = 1;
            out.push_back(K);
          }
        }
  }
  return out;
}

// Coarsening of complete families: the 8 children of a parent, all leaves
// and all flagged 2, become the parent, unless that breaks the 2:1 balance
// with the rest of the forest.  fl is carried over to the new
// leaves (0 for the parents).  A parent that the later refinement has to
// refine again for balance gets back the same 8 children, and the transfer
// (from the old mesh) keeps their values.
void coarsen_families(Forest &f, std::vector<char> &fl)
{
  std::unordered_map<uint64_t, int> count;
  for (size_t e = 0; e < f.leaves.size(); ++e) {
    const auto &L = f.leaves[e];
    if (fl[e] == 2 && L.level > 0) ++count[Forest::key(L.level - 1, L.ix >> 1, L.iy >> 1, L.iz >> 1)];
  }
  std::unordered_map<uint64_t, char> parent;  // candidate parents (1 = coarsen)
  for (auto &kv : count)
    if (kv.second == 8) parent[kv.first] = 1;
  if (parent.empty()) return;
  const std::vector<Forest::Leaf> old = f.leaves;
  const std::vector<char> oldfl = fl;
  for (int pass = 0; pass < 64; ++pass) {
    std::vector<Forest::Leaf> nl;
    std::vector<char> nf;
    std::unordered_map<uint64_t, char> added;
    for (size_t e = 0; e < old.size(); ++e) {
      const auto &L = old[e];
      const uint64_t pk = L.level > 0 ? Forest::key(L.level - 1, L.ix >> 1, L.iy >> 1, L.iz >> 1) : ~0ULL;
      auto it = parent.find(pk);
      if (L.level > 0 && it != parent.end() && it->second) {
        if (!added.count(pk)) {
          added[pk] = 1;
          nl.push_back({L.level - 1, L.ix >> 1, L.iy >> 1, L.iz >> 1});
          nf.push_back(0);
        }
      } else {
        nl.push_back(L);
        nf.push_back(oldfl[e] == 1 ? 1 : 0);
      }
    }
    // the flags follow their leaves into the Morton order of the new forest
    std::unordered_map<uint64_t, char> flk;
    for (size_t e = 0; e < nl.size(); ++e) flk[Forest::key(nl[e].level, nl[e].ix, nl[e].iy, nl[e].iz)] = nf[e];
    f.leaves.swap(nl);
    f.reindex();
    fl.assign(f.leaves.size(), 0);
    for (size_t e = 0; e < f.leaves.size(); ++e)
      fl[e] = flk[Forest::key(f.leaves[e].level, f.leaves[e].ix, f.leaves[e].iy, f.leaves[e].iz)];
    // undo the coarsenings that left a parent next to a leaf two levels finer
    bool undone = false;
    for (int K : too_coarse(f)) {
      const auto &L = f.leaves[K];
      auto it = parent.find(Forest::key(L.level, L.ix, L.iy, L.iz));
      if (it != parent.end() && it->second) {
        it->second = 0;
        undone = true;
      }
    }
    if (!undone) return;
  }
}

// Refinement of the leaves flagged 1, then of every leaf more than one level
// coarser than a neighbour, until the forest is 2:1 balanced.
int refine_balanced(Forest &f, std::vector<char> &fl, std::string &err)
{
  while (true) {
    bool any = false;
    std::vector<Forest::Leaf> nl;
    nl.reserve(f.leaves.size() + 8 * std::count(fl.begin(), fl.end(), 1));
    for (size_t e = 0; e < f.leaves.size(); ++e) {
      const Forest::Leaf L = f.leaves[e];
      if (fl[e] != 1) { nl.push_back(L); continue; }
      if (L.level >= Forest::LMAX) { err = "refinement deeper than the lattice allows"; return -1; }
      any = true;
      for (int c = 0; c < 8; ++c)
        nl.push_back({L.level + 1, 2 * L.ix + (c & 1), 2 * L.iy + ((c >> 1) & 1), 2 * L.iz + (c >> 2)});
    }
    if (!any) break;
    f.leaves.swap(nl);
    f.reindex();
    // 2:1 balance: a neighbour (26 directions) more than one level coarser
    // than a leaf is refined in the next pass
    fl.assign(f.leaves.size(), 0);
    for (int K : too_coarse(f)) fl[K] = 1;
  }
  return 0;
}

}  // namespace

int forest_refine(Forest &f, const std::vector<char> &flag, std::string &err)
{
  if (flag.size() != f.leaves.size()) { err = "one refinement flag per cell expected"; return -1; }
  for (char c : flag)
    if (c != 0 && c != 1 && c != 2) { err = "flags are 0 (keep), 1 (refine) or 2 (coarsen)"; return -1; }
  // refine_and_coarsen: the coarsening of complete families first, then the
  // refinement with 2:1 balance
  std::vector<char> fl(flag);
  coarsen_families(f, fl);
  return refine_balanced(f, fl, err);
}

// Lagrange basis of degree deg (1 or 2) on the equidistant nodes of [0,1]
double lagrange(int deg, int i, double t)
{
  if (deg == 1) return i ? t : 1.0 - t;
  if (i == 0) return (2.0 * (t - 0.5)) * (t - 1.0);
  if (i == 1) return (-4.0 * t) * (t - 1.0);
  return (2.0 * t) * (t - 0.5);
}

int forest_mesh(const Forest &f, Mesh &m, std::string &err, int deg)
{
  if (deg != 1 && deg != 2) { err = "forest meshes are Q1 or Q2"; return -1; }
  m = Mesh();
  const int n1 = deg + 1, ns = n1 * n1 * n1;
  m.deg = deg; m.n1 = n1; m.ns = ns;
  const int LM = Forest::LMAX;
  // node coordinates are integers in units of 1/deg of the finest lattice:
  // a leaf of level l spans deg * 2^(LM-l) units, its nodes sit 2^(LM-l) apart
  const int64_t NX = ((int64_t)f.nx << LM) * deg, NY = ((int64_t)f.ny << LM) * deg, NZ = ((int64_t)f.nz << LM) * deg;
  const double dl[3] = {(f.hi[0] - f.lo[0]) / (double)NX, (f.hi[1] - f.lo[1]) / (double)NY,
                        (f.hi[2] - f.lo[2]) / (double)NZ};
  const int nc = (int)f.leaves.size();
  if ((int64_t)nc * ns >= (1LL << 31)) { err = "mesh too large for 32-bit indices"; return -1; }
  m.n_cells = nc;
  // nodes, numbered in the (Morton) order of first appearance
  std::unordered_map<uint64_t, int> vid;
  vid.reserve((size_t)nc * (deg == 1 ? 2 : 9));
  std::vector<int64_t> vx;
  m.cell_nodes.resize((size_t)nc * ns);
  m.cell_h.resize((size_t)nc * 3);
  m.cell_lo.resize((size_t)nc * 3);
  m.cell_level.resize(nc);
  // a node on the + face of a periodic axis is the node of its - face
  const int64_t NN[3] = {NX, NY, NZ};
  const bool per[3] = {f.periodic(0), f.periodic(1), f.periodic(2)};
  auto canon = [&](int64_t &x, int64_t &y, int64_t &z) {
    int64_t *q[3] = {&x, &y, &z};
    for (int d = 0; d < 3; ++d)
      if (per[d]) *q[d] = ((*q[d] % NN[d]) + NN[d]) % NN[d];
  };
  for (int e = 0; e < nc; ++e) {
    const auto &L = f.leaves[e];
    const int64_t s = (int64_t)1 << (LM - L.level);
    const int64_t x0 = L.ix * s * deg, y0 = L.iy * s * deg, z0 = L.iz * s * deg;
    for (int a = 0; a < ns; ++a) {
      int64_t x = x0 + (a % n1) * s, y = y0 + ((a / n1) % n1) * s, z = z0 + (a / (n1 * n1)) * s;
      canon(x, y, z);
      auto it = vid.find(vkey(x, y, z));
      int v;
      if (it == vid.end()) {
        v = (int)(vx.size() / 3);
        vid.emplace(vkey(x, y, z), v);
        vx.push_back(x); vx.push_back(y); vx.push_back(z);
      } else {
        v = it->second;
      }
      m.cell_nodes[(size_t)e * ns + a] = v;
    }
    for (int d = 0; d < 3; ++d) m.cell_h[3 * e + d] = (double)(s * deg) * dl[d];
    m.cell_lo[3 * e + 0] = f.lo[0] + (double)x0 * dl[0];
    m.cell_lo[3 * e + 1] = f.lo[1] + (double)y0 * dl[1];
    m.cell_lo[3 * e + 2] = f.lo[2] + (double)z0 * dl[2];
    m.cell_level[e] = L.level;
  }
  const int nn = (int)(vx.size() / 3);
  if ((int64_t)nn * NC >= (1LL << 31)) { err = "mesh too large for 32-bit indices"; return -1; }
  m.n_nodes = nn;
  m.node_xyz.resize((size_t)nn * 3);
  for (int v = 0; v < nn; ++v)
    for (int d = 0; d < 3; ++d) m.node_xyz[3 * v + d] = f.lo[d] + (double)vx[3 * v + d] * dl[d];
  // hanging nodes: a node that lies on a face or an edge of a leaf touching
  // it without being one of that leaf's nodes follows the leaf's field there:
  // the tensor product of the degree-deg Lagrange values at its position
  // along the face or edge (Q1: 1/2 on an edge, 1/4 on a face; Q2: 3/8, 3/4,
  // -1/8 at a quarter of the edge, under 2:1 balance)
  std::vector<std::vector<std::pair<int, double>>> line(nn);
  std::vector<char> hanging(nn, 0);
  const int64_t NL[3] = {(int64_t)f.nx << LM, (int64_t)f.ny << LM, (int64_t)f.nz << LM};
  for (int v = 0; v < nn; ++v) {
    const int64_t P[3] = {vx[3 * v], vx[3 * v + 1], vx[3 * v + 2]};
    // finest-lattice cells whose closure holds the node: floor((P-1)/deg)
    // and floor(P/deg) in every direction; on a periodic axis the cell below
    // the - face is the last cell, and the node is seen one period up
    int64_t cand[3][2], shift[3][2];
    for (int d = 0; d < 3; ++d) {
      cand[d][0] = P[d] > 0 ? (P[d] - 1) / deg : -1;
      cand[d][1] = P[d] / deg;
      shift[d][0] = shift[d][1] = 0;
      if (per[d] && cand[d][0] < 0) { cand[d][0] = NL[d] - 1; shift[d][0] = NN[d]; }
    }
    for (int o = 0; o < 8 && !hanging[v]; ++o) {
      const int b[3] = {o & 1, (o >> 1) & 1, o >> 2};
      const int64_t cx = cand[0][b[0]], cy = cand[1][b[1]], cz = cand[2][b[2]];
      if (cx < 0 || cy < 0 || cz < 0 || cx >= NL[0] || cy >= NL[1] || cz >= NL[2]) continue;
      const int K = f.locate(LM, (int)cx, (int)cy, (int)cz);
      if (K < 0) { err = "forest lookup failed"; return -1; }
      const auto &L = f.leaves[K];
      const int64_t s = (int64_t)1 << (LM - L.level), ext = s * deg;
      const int64_t O[3] = {L.ix * ext, L.iy * ext, L.iz * ext};
      const int64_t r[3] = {P[0] + shift[0][b[0]] - O[0], P[1] + shift[1][b[1]] - O[1],
                            P[2] + shift[2][b[2]] - O[2]};
      bool node = true, bnd = false;
      for (int d = 0; d < 3; ++d) {
        if (r[d] < 0 || r[d] > ext) { err = "forest lookup off the node"; return -1; }
        if (r[d] % s) node = false;
        if (r[d] == 0 || r[d] == ext) bnd = true;
      }
      if (node) continue;
      if (!bnd) { err = "node inside a leaf"; return -1; }
      hanging[v] = 1;
      // free directions: along the face or edge (0 < r < ext)
      int fd[3], nf = 0;
      for (int d = 0; d < 3; ++d)
        if (r[d] > 0 && r[d] < ext) {
          if ((2 * r[d]) % s) { err = "unbalanced forest (hanging node off the half lattice)"; return -1; }
          fd[nf++] = d;
        }
      int cnt = 1;
      for (int k = 0; k < nf; ++k) cnt *= n1;
      for (int t = 0; t < cnt; ++t) {
        int64_t q[3] = {r[0], r[1], r[2]};
        double w = 1.0;
        int tt = t;
        for (int k = 0; k < nf; ++k) {
          const int i = tt % n1, d = fd[k];
          tt /= n1;
          q[d] = i * s;
          w *= lagrange(deg, i, (double)r[d] / (double)ext);
        }
        if (w == 0.0) continue;
        int64_t mx = O[0] + q[0], my = O[1] + q[1], mz = O[2] + q[2];
        canon(mx, my, mz);
        auto it = vid.find(vkey(mx, my, mz));
        if (it == vid.end()) { err = "master node missing"; return -1; }
        line[v].push_back({it->second, w});
      }
    }
  }
  // close the lines: masters that are hanging themselves are replaced by
  // their own lines until no hanging master remains
  for (int v = 0; v < nn; ++v) {
    if (!hanging[v]) continue;
    for (int pass = 0; pass < 32; ++pass) {
      bool changed = false;
      std::map<int, double> acc;
      for (auto &mw : line[v]) {
        if (hanging[mw.first]) {
          for (auto &mw2 : line[mw.first]) acc[mw2.first] += mw.second * mw2.second;
          changed = true;
        } else {
          acc[mw.first] += mw.second;
        }
      }
      line[v].assign(acc.begin(), acc.end());
      if (!changed) break;
      if (pass == 31) { err = "cyclic hanging-node constraints"; return -1; }
    }
  }
  m.hang_ptr.assign(nn + 1, 0);
  for (int v = 0; v < nn; ++v) {
    m.hang_ptr[v + 1] = m.hang_ptr[v] + (hanging[v] ? (int)line[v].size() : 0);
    if (hanging[v])
      for (auto &mw : line[v]) {
        m.hang_m.push_back(mw.first);
        m.hang_w.push_back(mw.second);
      }
  }
  // boundary faces and Dirichlet masks of the non-hanging nodes
  m.node_mask.assign(nn, 0u);
  for (int v = 0; v < nn; ++v) {
    if (hanging[v]) { m.node_mask[v] = HANG_BIT; continue; }
    const bool on[6] = {vx[3 * v] == 0, vx[3 * v] == NX, vx[3 * v + 1] == 0, vx[3 * v + 1] == NY,
                        vx[3 * v + 2] == 0, vx[3 * v + 2] == NZ};
    uint32_t mk = 0;
    for (int s = 0; s < 6; ++s)
      if (on[s]) mk |= dirichlet_mask(f.side_id[s]);
    m.node_mask[v] = mk;
  }
  for (int e = 0; e < nc; ++e) {
    const auto &L = f.leaves[e];
    const int64_t s = (int64_t)1 << (LM - L.level);
    const bool on[6] = {L.ix == 0, (L.ix + 1) * s == NL[0], L.iy == 0, (L.iy + 1) * s == NL[1],
                        L.iz == 0, (L.iz + 1) * s == NL[2]};
    for (int lf = 0; lf < 6; ++lf)
      if (on[lf] && f.side_id[lf] != 1) {
        m.faces.push_back(e);
        m.faces.push_back(lf);
        m.faces.push_back(f.side_id[lf]);
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

int forest_transfer_map(const Forest &fo, const Mesh &mo, const Forest &fn, const Mesh &mn, std::vector<int> &cell,
                        std::vector<float> &t, std::string &err)
{
  const int LM = Forest::LMAX;
  if (fo.nx != fn.nx || fo.ny != fn.ny || fo.nz != fn.nz) { err = "forests over different boxes"; return -1; }
  if ((int)fo.leaves.size() != mo.n_cells || (int)fn.leaves.size() != mn.n_cells) {
    err = "mesh does not belong to the forest";
    return -1;
  }
  const int deg = mn.deg;
  const int64_t NL[3] = {(int64_t)fn.nx << LM, (int64_t)fn.ny << LM, (int64_t)fn.nz << LM};
  // node coordinates in units of 1/deg of the finest lattice
  const double dl[3] = {(fn.hi[0] - fn.lo[0]) / (double)(NL[0] * deg), (fn.hi[1] - fn.lo[1]) / (double)(NL[1] * deg),
                        (fn.hi[2] - fn.lo[2]) / (double)(NL[2] * deg)};
  cell.resize(mn.n_nodes);
  t.resize((size_t)mn.n_nodes * 3);
  for (int v = 0; v < mn.n_nodes; ++v) {
    int64_t P[3], C[3];
    for (int d = 0; d < 3; ++d) {
      P[d] = (int64_t)std::llround((mn.node_xyz[3 * v + d] - fn.lo[d]) / dl[d]);
      C[d] = std::min(P[d] / deg, NL[d] - 1);
    }
    const int K = fo.locate(LM, (int)C[0], (int)C[1], (int)C[2]);
    if (K < 0) { err = "node outside the old forest"; return -1; }
    const auto &L = fo.leaves[K];
    const int64_t ext = ((int64_t)1 << (LM - L.level)) * deg;
    // k / (deg 2^(LM - level)) with k <= deg 2^LM: dyadic (Q1) or a quarter
    // step (Q2) of at most 14 bits, exact in float
    cell[v] = K;
    t[3 * (size_t)v] = (float)((double)(P[0] - L.ix * ext) / (double)ext);
    t[3 * (size_t)v + 1] = (float)((double)(P[1] - L.iy * ext) / (double)ext);
    t[3 * (size_t)v + 2] = (float)((double)(P[2] - L.iz * ext) / (double)ext);
  }
  return 0;
}

int forest_transfer(const Forest &fo, const Mesh &mo, const double *Uo, const Forest &fn, const Mesh &mn,
                    double *Un, std::string &err)
{
  if (mo.deg != mn.deg) { err = "old and new mesh of different degree"; return -1; }
  std::vector<int> cell;
  std::vector<float> tt;
  if (forest_transfer_map(fo, mo, fn, mn, cell, tt, err)) return -1;
  const int deg = mo.deg, n1 = deg + 1, ns = mo.ns;
  for (int v = 0; v < mn.n_nodes; ++v) {
    const double t[3] = {tt[3 * (size_t)v], tt[3 * (size_t)v + 1], tt[3 * (size_t)v + 2]};
    const int *cn = &mo.cell_nodes[(size_t)cell[v] * ns];
    for (int c = 0; c < NC; ++c) {
      double u = 0.0;
      for (int a = 0; a < ns; ++a) {
        const double w = lagrange(deg, a % n1, t[0]) * lagrange(deg, (a / n1) % n1, t[1]) *
                         lagrange(deg, a / (n1 * n1), t[2]);
        if (w != 0.0) u += w * Uo[(size_t)cn[a] * NC + c];
      }
      Un[(size_t)v * NC + c] = u;
    }
  }
  return 0;
}

}  // namespace glh
