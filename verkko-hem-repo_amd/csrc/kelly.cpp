// Kelly error estimator and fixed-number refinement flags for the Q1 and Q2
// forest meshes of amr.hpp.
//
// Restates what femgl/src/refine.cc:146-158 asks of deal.II:
// KellyErrorEstimator<3>::estimate(dof_handler, QGauss<2>(k+1), {} (no
// Neumann data), local_solution, estimated_error_per_cell) and
// parallel::distributed::GridRefinement::refine_and_coarsen_fixed_number
// (refine 0.3, coarsen 0).  The estimator, as documented by deal.II:
//   eta_K^2 = sum over faces F of K of  (h_K / 24) int_F sum_c [d u_c / dn]^2
// with h_K the cell diameter, [.] the jump across interior faces (0 on the
// boundary: no Neumann data), all 18 components, by the (k+1)^2-point Gauss
// rule on the face; faces between cells of different levels are integrated
// on the smaller face.  eta is stored in
// float, as the reference's Vector<float>.
#include <algorithm>
#include <cmath>
#include <numeric>

#include "amr.hpp"

namespace glh {

namespace {

double dlagrange(int deg, int i, double t)
{
  if (deg == 1) return i ? 1.0 : -1.0;
  if (i == 0) return 4.0 * t - 3.0;
  if (i == 1) return 4.0 - 8.0 * t;
  return 4.0 * t - 1.0;
}

// gradient (3 x 18) of the Q_deg field of leaf e at reference point xi
void field_grad(const Mesh &m, const double *U, int e, const double xi[3], double g[3][NC])
{
  const double *h = &m.cell_h[3 * e];
  const int deg = m.deg, n1 = m.n1;
  for (int k = 0; k < 3; ++k)
    for (int c = 0; c < NC; ++c) g[k][c] = 0.0;
  for (int a = 0; a < m.ns; ++a) {
    const int ax = a % n1, ay = (a / n1) % n1, az = a / (n1 * n1);
    const double lx = lagrange(deg, ax, xi[0]), ly = lagrange(deg, ay, xi[1]), lz = lagrange(deg, az, xi[2]);
    const double dx = dlagrange(deg, ax, xi[0]) * ly * lz / h[0];
    const double dy = lx * dlagrange(deg, ay, xi[1]) * lz / h[1];
    const double dz = lx * ly * dlagrange(deg, az, xi[2]) / h[2];
    const double *u = U + (size_t)m.cell_nodes[(size_t)e * m.ns + a] * NC;
    for (int c = 0; c < NC; ++c) {
      g[0][c] += dx * u[c];
      g[1][c] += dy * u[c];
      g[2][c] += dz * u[c];
    }
  }
}

}  // namespace

int kelly_estimate(const Forest &f, const Mesh &m, const double *U, float *eta, std::string &err)
{
  const int nc = (int)f.leaves.size();
  if (m.n_cells != nc || (m.deg != 1 && m.deg != 2)) {
    err = "Kelly estimator: Q1 or Q2 mesh of this forest expected";
    return -1;
  }
  std::vector<double> e2(nc, 0.0);
  // QGauss<2>(deg + 1) on the face, as refine.cc:148-151 asks
  const int ng = m.deg + 1;
  const double gq2[2] = {0.5 - 0.5 / std::sqrt(3.0), 0.5 + 0.5 / std::sqrt(3.0)}, gw2[2] = {0.5, 0.5};
  const double gq3[3] = {0.5 - 0.5 * std::sqrt(0.6), 0.5, 0.5 + 0.5 * std::sqrt(0.6)},
               gw3[3] = {5.0 / 18.0, 8.0 / 18.0, 5.0 / 18.0};
  const double *gq = ng == 2 ? gq2 : gq3, *gw = ng == 2 ? gw2 : gw3;
  for (int e = 0; e < nc; ++e) {
    const auto &L = f.leaves[e];
    const int l = L.level;
    const int n[3] = {f.nx << l, f.ny << l, f.nz << l};
    const int idx[3] = {L.ix, L.iy, L.iz};
    for (int d = 0; d < 3; ++d)
      for (int sg = 0; sg < 2; ++sg) {
        int nb[3] = {idx[0], idx[1], idx[2]};
        nb[d] += sg ? 1 : -1;
        // across a periodic face the neighbour sits one period away
        double wrap = 0.0;
        if (nb[d] < 0 || nb[d] >= n[d]) {
          if (!f.periodic(d)) continue;  // boundary face: no Neumann data, no jump
          wrap = (sg ? 1.0 : -1.0) * (f.hi[d] - f.lo[d]);
          nb[d] = (nb[d] + n[d]) % n[d];
        }
        const int N = f.locate(l, nb[0], nb[1], nb[2]);
        if (N < 0) continue;                         // finer neighbours integrate their pieces
        const int ln = f.leaves[N].level;
        if (ln == l && sg == 0) continue;            // same level: integrated from the lower cell
        // the piece is this leaf's face; 2x2 Gauss points on it
        const double *h = &m.cell_h[3 * e];
        const int t1 = (d + 1) % 3, t2 = (d + 2) % 3;
        const double area = h[t1] * h[t2];
        double s = 0.0;
        for (int q1 = 0; q1 < ng; ++q1)
          for (int q2 = 0; q2 < ng; ++q2) {
            double P[3];
            P[d] = m.cell_lo[3 * e + d] + (sg ? h[d] : 0.0);
            P[t1] = m.cell_lo[3 * e + t1] + gq[q1] * h[t1];
            P[t2] = m.cell_lo[3 * e + t2] + gq[q2] * h[t2];
            double xa[3], xb[3];
            for (int k = 0; k < 3; ++k) {
              xa[k] = (P[k] - m.cell_lo[3 * e + k]) / h[k];
              xb[k] = (P[k] - (k == d ? wrap : 0.0) - m.cell_lo[3 * N + k]) / m.cell_h[3 * N + k];
            }
            xa[d] = sg ? 1.0 : 0.0;
            xb[d] = sg ? 0.0 : 1.0;
            double ga[3][NC], gb[3][NC];
            field_grad(m, U, e, xa, ga);
            field_grad(m, U, N, xb, gb);
            double j2 = 0.0;
            for (int c = 0; c < NC; ++c) {
              const double jmp = ga[d][c] - gb[d][c];
              j2 += jmp * jmp;
            }
            s += gw[q1] * gw[q2] * area * j2;
          }
        auto diam = [&](int k) {
          const double *hk = &m.cell_h[3 * k];
          return std::sqrt(hk[0] * hk[0] + hk[1] * hk[1] + hk[2] * hk[2]);
        };
        e2[e] += diam(e) / 24.0 * s;
        e2[N] += diam(N) / 24.0 * s;
      }
  }
  for (int e = 0; e < nc; ++e) eta[e] = (float)std::sqrt(e2[e]);
  return 0;
}

void flag_refine_coarsen(int n, const float *eta, double refine, double coarsen, std::vector<char> &flag)
{
  flag.assign(n, 0);
  const int kr = std::max(0, std::min(n, (int)std::floor(refine * n)));
  const int kc = std::max(0, std::min(n - kr, (int)std::floor(coarsen * n)));
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return eta[a] > eta[b]; });
  for (int i = 0; i < kr; ++i) flag[ord[i]] = 1;
  // smallest indicators: the end of the same order (ties in reverse cell order)
  for (int i = 0; i < kc; ++i) flag[ord[n - 1 - i]] = 2;
}

void flag_fixed_number(int n, const float *eta, double frac, std::vector<char> &flag)
{
  flag.assign(n, 0);
  const int k = std::max(0, std::min(n, (int)std::floor(frac * n)));
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  // largest indicators first; equal values in cell order
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return eta[a] > eta[b]; });
  for (int i = 0; i < k; ++i) flag[ord[i]] = 1;
}

}  // namespace glh
