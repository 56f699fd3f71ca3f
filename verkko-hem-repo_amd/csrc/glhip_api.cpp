// C ABI of the MI355X Ginzburg-Landau Newton-step library: context, per-cycle
// upload, assembly, GMRES + point-block Jacobi, damped Newton line search and
// free energy.  Host orchestration of the HIP kernels in gl_assemble.hip and
// gl_linalg.hip.  See include/glhip.h for the reference methods each entry
// point replaces.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "glhip.h"
#include "glhip_dev.hpp"
#include "mesh.hpp"

using glh::NC;

struct glh_mesh {
  glh::Mesh m;
};

namespace {

template <typename T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  hipError_t alloc(size_t count)
  {
    if (count <= n && p) return hipSuccess;
    release();
    hipError_t e = hipMalloc((void **)&p, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) { p = nullptr; return e; }
    n = count;
    return hipSuccess;
  }
  void release()
  {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

}  // namespace

struct glh_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  bool have_mat = false, have_mesh = false;
  glh::Material mat{};
  // mesh
  int deg = 1, ns = 8, nq = 8, npair = 36;
  int n_cells = 0, n_nodes = 0, nnzb = 0;
  std::vector<int> color_ptr;
  // reference tables
  DBuf<double> tN, tdN, tw, tWp, tSp, tMf;
  DBuf<int> tpa, tpb;
  // mesh arrays
  DBuf<int> cell_nodes, row_ptr, col, diag, cell_blk, color_cells;
  DBuf<double> cell_h;
  DBuf<uint32_t> cell_face, node_mask;
  // state vectors (FP64)
  DBuf<double> U, U0, dU, X, rhs, res, Fcell, red_partial, red_out;
  // Jacobian and preconditioner
  DBuf<double> J64, D64;
  DBuf<float> J32, D32;
  bool j_fp32 = false, j_valid = false, d_valid = false;
  DBuf<int> status;
  // GMRES workspace
  DBuf<double> gV64, gZ64, gW64, gh64, gp64, gy64;
  DBuf<float> gV32, gZ32, gW32, gh32, gp32, gy32, gb32, gx32;
  DBuf<double> ir_r, ir_c;
  // timing
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double asm_ms = 0.0, asm_launches = 0;
  // live kernel timing (glh_set_timing): event pairs around each SpMV
  // launch on the context stream, read back when the pool fills or a solve ends
  bool timing = false;
  long reorth_count = 0;
  std::vector<hipEvent_t> ev_pool;
  int ev_used = 0;
  double spmv_ms_total = 0.0, spmv_count = 0;
  double jasm_ms_total = 0.0, jasm_count = 0, rasm_ms_total = 0.0, rasm_count = 0;
  double last_rhs_norm = -1.0;

  glh::RefTables tables() const
  {
    return glh::RefTables{tN.p, tdN.p, tw.p, tWp.p, tSp.p, tMf.p, tpa.p, tpb.p};
  }
  glh::MeshDev meshdev() const
  {
    return glh::MeshDev{cell_nodes.p, cell_h.p, cell_face.p, node_mask.p, cell_blk.p};
  }
};

namespace {

int fail(glh_ctx *c, int code, const std::string &msg)
{
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(ctx, ...)                                                                       \
  do {                                                                                         \
    hipError_t e_ = (__VA_ARGS__);                                                             \
    if (e_ != hipSuccess)                                                                      \
      return fail((ctx), GLH_ERR_HIP, std::string(#__VA_ARGS__) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
hipError_t upload(DBuf<T> &b, const T *src, size_t n, hipStream_t s)
{
  hipError_t e = b.alloc(n);
  if (e != hipSuccess) return e;
  if (n == 0) return hipSuccess;
  return hipMemcpyAsync(b.p, src, n * sizeof(T), hipMemcpyHostToDevice, s);
}

// 1D Gauss points on [0,1] and Lagrange basis on equispaced support points
void gauss01(int n, double *x, double *w)
{
  if (n == 2) {
    const double d = 0.5 / std::sqrt(3.0);
    x[0] = 0.5 - d; x[1] = 0.5 + d; w[0] = w[1] = 0.5;
  } else {
    const double d = 0.5 * std::sqrt(0.6);
    x[0] = 0.5 - d; x[1] = 0.5; x[2] = 0.5 + d;
    w[0] = w[2] = 5.0 / 18.0; w[1] = 8.0 / 18.0;
  }
}
void lagrange(int deg, double x, double *L, double *dL)
{
  if (deg == 1) {
    L[0] = 1 - x; L[1] = x; dL[0] = -1; dL[1] = 1;
  } else {
    L[0] = 2 * (x - 0.5) * (x - 1); L[1] = -4 * x * (x - 1); L[2] = 2 * x * (x - 0.5);
    dL[0] = 4 * x - 3; dL[1] = 4 - 8 * x; dL[2] = 4 * x - 1;
  }
}

int build_tables(glh_ctx *c, int deg)
{
  const int n1 = deg + 1, ns = n1 * n1 * n1, nq = ns, npair = ns * (ns + 1) / 2;
  double gx[3], gw[3], L[3][3], dL[3][3];
  gauss01(n1, gx, gw);
  for (int i = 0; i < n1; ++i) lagrange(deg, gx[i], L[i], dL[i]);
  std::vector<double> N(nq * ns), dN(nq * ns * 3), w(nq);
  for (int q = 0; q < nq; ++q) {
    const int qx = q % n1, qy = (q / n1) % n1, qz = q / (n1 * n1);
    w[q] = gw[qx] * gw[qy] * gw[qz];
    for (int a = 0; a < ns; ++a) {
      const int ax = a % n1, ay = (a / n1) % n1, az = a / (n1 * n1);
      N[q * ns + a] = L[qx][ax] * L[qy][ay] * L[qz][az];
      dN[(q * ns + a) * 3 + 0] = dL[qx][ax] * L[qy][ay] * L[qz][az];
      dN[(q * ns + a) * 3 + 1] = L[qx][ax] * dL[qy][ay] * L[qz][az];
      dN[(q * ns + a) * 3 + 2] = L[qx][ax] * L[qy][ay] * dL[qz][az];
    }
  }
  std::vector<double> Wp((size_t)npair * nq), Sp((size_t)npair * 9);
  std::vector<int> pa(npair), pb(npair);
  int p = 0;
  for (int a = 0; a < ns; ++a)
    for (int b = a; b < ns; ++b, ++p) {
      pa[p] = a; pb[p] = b;
      for (int q = 0; q < nq; ++q) Wp[(size_t)p * nq + q] = w[q] * N[q * ns + a] * N[q * ns + b];
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) {
          double s = 0;
          for (int q = 0; q < nq; ++q) s += w[q] * dN[(q * ns + a) * 3 + k] * dN[(q * ns + b) * 3 + l];
          Sp[(size_t)p * 9 + k * 3 + l] = s;
        }
    }
  // face mass on each unit face: QGauss<2>(k+1) (assemble.cc:124)
  std::vector<double> Mf((size_t)6 * ns * ns, 0.0);
  for (int f = 0; f < 6; ++f) {
    const int dn = f / 2;
    int t0 = (dn + 1) % 3, t1 = (dn + 2) % 3;
    if (t0 > t1) std::swap(t0, t1);
    for (int i0 = 0; i0 < n1; ++i0)
      for (int i1 = 0; i1 < n1; ++i1) {
        double xi[3];
        xi[dn] = (f & 1) ? 1.0 : 0.0;
        xi[t0] = gx[i0];
        xi[t1] = gx[i1];
        const double wf = gw[i0] * gw[i1];
        double Lx[3][3], dLx[3][3];
        for (int k = 0; k < 3; ++k) lagrange(deg, xi[k], Lx[k], dLx[k]);
        std::vector<double> Ns(ns);
        for (int a = 0; a < ns; ++a) Ns[a] = Lx[0][a % n1] * Lx[1][(a / n1) % n1] * Lx[2][a / (n1 * n1)];
        for (int a = 0; a < ns; ++a)
          for (int b = 0; b < ns; ++b) Mf[((size_t)f * ns + a) * ns + b] += wf * Ns[a] * Ns[b];
      }
  }
  hipStream_t s = c->stream;
  HIPCHK(c, upload(c->tN, N.data(), N.size(), s));
  HIPCHK(c, upload(c->tdN, dN.data(), dN.size(), s));
  HIPCHK(c, upload(c->tw, w.data(), w.size(), s));
  HIPCHK(c, upload(c->tWp, Wp.data(), Wp.size(), s));
  HIPCHK(c, upload(c->tSp, Sp.data(), Sp.size(), s));
  HIPCHK(c, upload(c->tMf, Mf.data(), Mf.size(), s));
  HIPCHK(c, upload(c->tpa, pa.data(), pa.size(), s));
  HIPCHK(c, upload(c->tpb, pb.data(), pb.size(), s));
  HIPCHK(c, hipStreamSynchronize(s));
  c->deg = deg; c->ns = ns; c->nq = nq; c->npair = npair;
  return GLH_OK;
}

int check_ready(glh_ctx *c, bool need_mat = true)
{
  if (!c) return GLH_ERR_ARG;
  if (!c->have_mesh) return fail(c, GLH_ERR_STATE, "no mesh uploaded");
  if (need_mat && !c->have_mat) return fail(c, GLH_ERR_STATE, "material not set");
  HIPCHK(c, hipSetDevice(c->device));
  return GLH_OK;
}

size_t ndofs(const glh_ctx *c) { return (size_t)c->n_nodes * NC; }

// ||v|| of an FP64 device vector (deterministic), result on host
int norm64(glh_ctx *c, const double *v, double *out)
{
  const size_t n = ndofs(c);
  HIPCHK(c, c->red_partial.alloc(glh::MDOT_GRID * 2));
  HIPCHK(c, c->red_out.alloc(4));
  HIPCHK(c, glh::launch_mdot<double>(n, 1, v, n, v, c->red_partial.p, c->red_out.p, false, c->stream));
  double h = 0;
  HIPCHK(c, hipMemcpyAsync(&h, c->red_out.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *out = std::sqrt(h);
  return GLH_OK;
}

// assemble (J and/or -R) at state Ustate into (J, rhs_out)
int assemble_at(glh_ctx *c, const double *Ustate, bool do_j, bool do_r, bool fp32, double *rhs_out, bool timed)
{
  void *J = nullptr;
  if (do_j) {
    const size_t nv = (size_t)c->nnzb * 324;
    if (fp32) { HIPCHK(c, c->J32.alloc(nv)); J = c->J32.p; }
    else { HIPCHK(c, c->J64.alloc(nv)); J = c->J64.p; }
  }
  if (do_r) HIPCHK(c, hipMemsetAsync(rhs_out, 0, ndofs(c) * sizeof(double), c->stream));
  const glh::RefTables T = c->tables();
  const glh::MeshDev M = c->meshdev();
  timed = timed || c->timing;
  if (timed) HIPCHK(c, hipEventRecord(c->ev0, c->stream));
  const int ncol = (int)c->color_ptr.size() - 1;
  for (int k = 0; k < ncol; ++k) {
    const int off = c->color_ptr[k], n = c->color_ptr[k + 1] - off;
    HIPCHK(c, glh::launch_assemble(c->deg, do_j, do_r, fp32, T, M, c->mat, Ustate, c->color_cells.p + off, n, J,
                                   rhs_out, c->stream));
  }
  if (timed) {
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->asm_ms = ms;
    c->asm_launches = ncol;
    if (c->timing) {
      if (do_j) { c->jasm_ms_total += ms; c->jasm_count += 1; }
      else { c->rasm_ms_total += ms; c->rasm_count += 1; }
    }
  }
  if (do_j) {
    c->j_fp32 = fp32;
    c->j_valid = true;
    c->d_valid = false;
  }
  return GLH_OK;
}

int ensure_precond(glh_ctx *c)
{
  if (c->d_valid) return GLH_OK;
  HIPCHK(c, c->status.alloc(1));
  HIPCHK(c, hipMemsetAsync(c->status.p, 0, sizeof(int), c->stream));
  const size_t nv = (size_t)c->n_nodes * 324;
  if (c->j_fp32) {
    HIPCHK(c, c->D32.alloc(nv));
    HIPCHK(c, glh::launch_bj_invert<float>(c->n_nodes, c->diag.p, c->J32.p, c->D32.p, c->status.p, c->stream));
  } else {
    HIPCHK(c, c->D64.alloc(nv));
    HIPCHK(c, glh::launch_bj_invert<double>(c->n_nodes, c->diag.p, c->J64.p, c->D64.p, c->status.p, c->stream));
  }
  int st = 0;
  HIPCHK(c, hipMemcpyAsync(&st, c->status.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (st) return fail(c, GLH_ERR_SINGULAR, "singular 18x18 diagonal block in point-block Jacobi");
  c->d_valid = true;
  return GLH_OK;
}

template <typename TV>
struct Work;
template <>
struct Work<double> {
  static DBuf<double> &V(glh_ctx *c) { return c->gV64; }
  static DBuf<double> &Z(glh_ctx *c) { return c->gZ64; }
  static DBuf<double> &W(glh_ctx *c) { return c->gW64; }
  static DBuf<double> &h(glh_ctx *c) { return c->gh64; }
  static DBuf<double> &part(glh_ctx *c) { return c->gp64; }
  static DBuf<double> &y(glh_ctx *c) { return c->gy64; }
};
template <>
struct Work<float> {
  static DBuf<float> &V(glh_ctx *c) { return c->gV32; }
  static DBuf<float> &Z(glh_ctx *c) { return c->gZ32; }
  static DBuf<float> &W(glh_ctx *c) { return c->gW32; }
  static DBuf<float> &h(glh_ctx *c) { return c->gh32; }
  static DBuf<float> &part(glh_ctx *c) { return c->gp32; }
  static DBuf<float> &y(glh_ctx *c) { return c->gy32; }
};

int flush_events(glh_ctx *c)
{
  if (c->ev_used == 0) return GLH_OK;
  HIPCHK(c, hipEventSynchronize(c->ev_pool[2 * c->ev_used - 1]));
  for (int i = 0; i < c->ev_used; ++i) {
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev_pool[2 * i], c->ev_pool[2 * i + 1]));
    c->spmv_ms_total += ms;
    c->spmv_count += 1;
  }
  c->ev_used = 0;
  return GLH_OK;
}

template <typename TJ, typename TV>
int spmv(glh_ctx *c, const TJ *J, const TV *x, TV *y)
{
  const bool t = c->timing && !c->ev_pool.empty();
  if (t) {
    if (2 * c->ev_used >= (int)c->ev_pool.size()) {
      int rc = flush_events(c);
      if (rc) return rc;
    }
    HIPCHK(c, hipEventRecord(c->ev_pool[2 * c->ev_used], c->stream));
  }
  HIPCHK(c, glh::launch_spmv<TJ, TV>(c->n_nodes, c->row_ptr.p, c->col.p, J, x, y, c->stream));
  if (t) {
    HIPCHK(c, hipEventRecord(c->ev_pool[2 * c->ev_used + 1], c->stream));
    ++c->ev_used;
  }
  return GLH_OK;
}

// Right-preconditioned restarted GMRES(m) with 18x18 point-block Jacobi and
// classical Gram-Schmidt with selective reorthogonalisation on the device.  Solves J x = b (b, x device vectors
// of type TV) to ||b - J x|| <= rtol ||b||.  Hessenberg/Givens on the host.
template <typename TJ, typename TV>
int gmres(glh_ctx *c, const TJ *J, const TJ *Dinv, const TV *b, TV *x, double rtol, int maxit, int m, int *its_out,
          double *rel_out)
{
  const size_t n = ndofs(c);
  hipStream_t s = c->stream;
  auto &V = Work<TV>::V(c);
  auto &Z = Work<TV>::Z(c);
  auto &W = Work<TV>::W(c);
  auto &hb = Work<TV>::h(c);
  auto &pt = Work<TV>::part(c);
  auto &yb = Work<TV>::y(c);
  HIPCHK(c, V.alloc(n * (m + 1)));
  HIPCHK(c, Z.alloc(n));
  HIPCHK(c, W.alloc(n));
  HIPCHK(c, hb.alloc(2 * (m + 2)));
  HIPCHK(c, V.alloc(n * (m + 1)));
  HIPCHK(c, pt.alloc((size_t)glh::MDOT_GRID * (m + 2)));
  HIPCHK(c, yb.alloc(m + 1));
  TV *hsum = hb.p, *hpass = hb.p + (m + 2);
  std::vector<TV> hh(m + 3);
  std::vector<double> H((size_t)(m + 1) * m, 0.0), cs(m), sn(m), g(m + 1), y(m);
  std::vector<TV> yh(m);
  HIPCHK(c, hipMemsetAsync(x, 0, n * sizeof(TV), s));
  // ||b||
  HIPCHK(c, glh::launch_mdot<TV>(n, 1, b, n, b, pt.p, hsum, false, s));
  TV bn2 = 0;
  HIPCHK(c, hipMemcpyAsync(&bn2, hsum, sizeof(TV), hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  const double bn = std::sqrt((double)bn2);
  const double target = rtol * bn;
  int it = 0;
  double res = bn;
  bool conv = false;
  while (true) {
    // V0 = b - J x
    if (it == 0) {
      HIPCHK(c, hipMemcpyAsync(V.p, b, n * sizeof(TV), hipMemcpyDeviceToDevice, s));
    } else {
      int rc = spmv<TJ, TV>(c, J, x, W.p);
      if (rc) return rc;
      HIPCHK(c, hipMemcpyAsync(V.p, b, n * sizeof(TV), hipMemcpyDeviceToDevice, s));
      HIPCHK(c, glh::launch_set_scalar<TV>(hpass, (TV)1, s));
      HIPCHK(c, glh::launch_sub_lincomb<TV>(n, 1, W.p, n, hpass, V.p, s));
    }
    HIPCHK(c, glh::launch_mdot<TV>(n, 1, V.p, n, V.p, pt.p, hsum, false, s));
    TV beta2 = 0;
    HIPCHK(c, hipMemcpyAsync(&beta2, hsum, sizeof(TV), hipMemcpyDeviceToHost, s));
    HIPCHK(c, hipStreamSynchronize(s));
    const double beta = std::sqrt((double)beta2);
    res = beta;
    if (beta <= target) { conv = true; break; }
    if (it >= maxit) break;
    HIPCHK(c, glh::launch_set_scalar<TV>(hpass, (TV)beta, s));
    HIPCHK(c, glh::launch_scale_by<TV>(n, V.p, hpass, V.p, s));
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int j = 0;
    bool est = false;
    while (j < m && it < maxit && !est) {
      TV *vj = V.p + n * j;
      TV *wv = V.p + n * (j + 1);  // the new direction is built in place
      HIPCHK(c, glh::launch_bj_apply<TJ, TV>(c->n_nodes, Dinv, vj, Z.p, s));
      int rc = spmv<TJ, TV>(c, J, Z.p, wv);
      if (rc) return rc;
      // classical Gram-Schmidt: h = V^T w together with ||w||^2 in one pass,
      // then w -= V h with ||w'||^2; a second pass only when
      // ||w'|| < ||w|| / sqrt(2) (Daniel-Gragg-Kaufman-Stewart criterion)
      HIPCHK(c, glh::launch_mdot<TV>(n, j + 2, V.p, n, wv, pt.p, hsum, false, s));
      HIPCHK(c, glh::launch_sub_lincomb_norm<TV>(n, j + 1, V.p, n, hsum, wv, pt.p, hpass + (j + 1), s));
      HIPCHK(c, hipMemcpyAsync(hh.data(), hsum, (j + 2) * sizeof(TV), hipMemcpyDeviceToHost, s));
      TV hn2 = 0;
      HIPCHK(c, hipMemcpyAsync(&hn2, hpass + (j + 1), sizeof(TV), hipMemcpyDeviceToHost, s));
      HIPCHK(c, hipStreamSynchronize(s));
      std::vector<double> h2(j + 1, 0.0);
      if ((double)hn2 < 0.5 * (double)hh[j + 1]) {
        HIPCHK(c, glh::launch_mdot<TV>(n, j + 1, V.p, n, wv, pt.p, hpass, false, s));
        HIPCHK(c, glh::launch_sub_lincomb_norm<TV>(n, j + 1, V.p, n, hpass, wv, pt.p, hpass + (j + 1), s));
        std::vector<TV> hp(j + 2);
        HIPCHK(c, hipMemcpyAsync(hp.data(), hpass, (j + 2) * sizeof(TV), hipMemcpyDeviceToHost, s));
        HIPCHK(c, hipStreamSynchronize(s));
        for (int i = 0; i <= j; ++i) h2[i] = (double)hp[i];
        hn2 = hp[j + 1];
        ++c->reorth_count;
      }
      const double hn = std::sqrt(std::max(0.0, (double)hn2));
      if (hn > 0) {
        HIPCHK(c, glh::launch_set_scalar<TV>(hpass, (TV)hn, s));
        HIPCHK(c, glh::launch_scale_by<TV>(n, wv, hpass, wv, s));
      }
      for (int i = 0; i <= j; ++i) H[(size_t)i * m + j] = (double)hh[i] + h2[i];
      for (int i = 0; i < j; ++i) {
        const double a = H[(size_t)i * m + j], b2 = H[(size_t)(i + 1) * m + j];
        H[(size_t)i * m + j] = cs[i] * a + sn[i] * b2;
        H[(size_t)(i + 1) * m + j] = -sn[i] * a + cs[i] * b2;
      }
      const double a = H[(size_t)j * m + j], rr = std::hypot(a, hn);
      cs[j] = rr > 0 ? a / rr : 1.0;
      sn[j] = rr > 0 ? hn / rr : 0.0;
      H[(size_t)j * m + j] = rr;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      ++j;
      ++it;
      if (std::fabs(g[j]) <= target || hn == 0) est = true;
    }
    for (int i = j - 1; i >= 0; --i) {
      double sum = g[i];
      for (int k = i + 1; k < j; ++k) sum -= H[(size_t)i * m + k] * y[k];
      y[i] = sum / H[(size_t)i * m + i];
    }
    for (int i = 0; i < j; ++i) yh[i] = (TV)y[i];
    HIPCHK(c, hipMemcpyAsync(yb.p, yh.data(), j * sizeof(TV), hipMemcpyHostToDevice, s));
    HIPCHK(c, glh::launch_lincomb<TV>(n, j, V.p, n, yb.p, W.p, s));
    HIPCHK(c, glh::launch_bj_apply<TJ, TV>(c->n_nodes, Dinv, W.p, Z.p, s));
    HIPCHK(c, glh::launch_axpy<TV, TV>(n, 1.0, Z.p, x, s));
  }
  {
    int rc = flush_events(c);
    if (rc) return rc;
  }
  *its_out = it;
  *rel_out = bn > 0 ? res / bn : 0.0;
  return conv ? GLH_OK : fail(c, GLH_ERR_NOCONV, "GMRES reached the maximum iteration number");
}

int solve_impl(glh_ctx *c, double rtol, int maxit, int restart, double stats[2])
{
  if (!c->j_valid) return fail(c, GLH_ERR_STATE, "no Jacobian assembled");
  if (restart < 1 || maxit < 1 || !(rtol > 0)) return fail(c, GLH_ERR_ARG, "bad solver parameters");
  int rc = ensure_precond(c);
  if (rc) return rc;
  const size_t n = ndofs(c);
  hipStream_t s = c->stream;
  int its = 0;
  double rel = 0;
  if (!c->j_fp32) {
    rc = gmres<double, double>(c, c->J64.p, c->D64.p, c->rhs.p, c->dU.p, rtol, maxit, restart, &its, &rel);
  } else {
    // FP32 GMRES inside an FP64 iterative-refinement loop:
    //   r = rhs - J32 dU (FP64 arithmetic), solve J32 c = r in FP32, dU += c
    HIPCHK(c, c->gb32.alloc(n));
    HIPCHK(c, c->gx32.alloc(n));
    HIPCHK(c, c->ir_r.alloc(n));
    HIPCHK(c, hipMemsetAsync(c->dU.p, 0, n * sizeof(double), s));
    double bn = 0;
    rc = norm64(c, c->rhs.p, &bn);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(c->ir_r.p, c->rhs.p, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    double rn = bn;
    const double inner = std::max(rtol, 1e-4);
    for (int sweep = 0; sweep < 20 && rn > rtol * bn && its < maxit; ++sweep) {
      HIPCHK(c, glh::launch_convert<double, float>(n, c->ir_r.p, c->gb32.p, s));
      int it2 = 0;
      double rel2 = 0;
      rc = gmres<float, float>(c, c->J32.p, c->D32.p, c->gb32.p, c->gx32.p, std::max(inner, rtol * bn / rn),
                               maxit - its, restart, &it2, &rel2);
      its += it2;
      if (rc && rc != GLH_ERR_NOCONV) return rc;
      HIPCHK(c, glh::launch_axpy<float, double>(n, 1.0, c->gx32.p, c->dU.p, s));
      HIPCHK(c, glh::launch_residual_spmv<float>(c->n_nodes, c->row_ptr.p, c->col.p, c->J32.p, c->dU.p, c->rhs.p,
                                                 c->ir_r.p, s));
      rc = norm64(c, c->ir_r.p, &rn);
      if (rc) return rc;
    }
    rel = bn > 0 ? rn / bn : 0;
    rc = rel <= rtol ? GLH_OK : fail(c, GLH_ERR_NOCONV, "mixed-precision GMRES did not reach the tolerance");
  }
  if (stats) { stats[0] = its; stats[1] = rel; }
  if (rc) return rc;
  // constraints_newton_update.distribute (solve.cc:181): Dirichlet components to zero
  HIPCHK(c, glh::launch_mask_zero(c->n_nodes, c->node_mask.p, c->dU.p, s));
  return GLH_OK;
}

int newton_step_impl(glh_ctx *c, double step, int damped, double out[4])
{
  double prev = 0;
  int rc = norm64(c, c->rhs.p, &prev);  // previous_residual = ||system_rhs|| (iteration.cc:130)
  if (rc) return rc;
  const int n_try = damped ? 100 : 1;
  double a = 1.0, cur = 0.0;
  int i = 0;
  for (i = 0; i < n_try; ++i) {
    a = damped ? std::pow(step, (double)i) : 1.0;
    HIPCHK(c, glh::launch_trial(c->n_nodes, c->U.p, c->dU.p, a, c->node_mask.p, c->X.p, c->stream));
    rc = assemble_at(c, c->X.p, false, true, false, c->res.p, false);
    if (rc) return rc;
    rc = norm64(c, c->res.p, &cur);
    if (rc) return rc;
    if (cur < prev) break;
  }
  if (i == n_try) i = n_try - 1;
  HIPCHK(c, hipMemcpyAsync(c->U.p, c->X.p, ndofs(c) * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (out) { out[0] = prev; out[1] = cur; out[2] = a; out[3] = i + 1; }
  return GLH_OK;
}

}  // namespace

// ===================================================================== ABI
extern "C" {

int glh_create(int device, glh_ctx **out)
{
  if (!out) return GLH_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return GLH_ERR_HIP;
  if (device < 0 || device >= ndev) return GLH_ERR_ARG;
  glh_ctx *c = new glh_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    delete c;
    return GLH_ERR_HIP;
  }
  *out = c;
  return GLH_OK;
}

int glh_destroy(glh_ctx *c)
{
  if (!c) return GLH_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  DBuf<double> *dbs[] = {&c->tN, &c->tdN, &c->tw, &c->tWp, &c->tSp, &c->tMf, &c->cell_h, &c->U, &c->U0, &c->dU, &c->X,
                         &c->rhs, &c->res, &c->Fcell, &c->red_partial, &c->red_out, &c->J64, &c->D64, &c->gV64,
                         &c->gZ64, &c->gW64, &c->gh64, &c->gp64, &c->gy64, &c->ir_r, &c->ir_c};
  for (auto *b : dbs) b->release();
  DBuf<float> *fbs[] = {&c->J32, &c->D32, &c->gV32, &c->gZ32, &c->gW32, &c->gh32, &c->gp32, &c->gy32, &c->gb32,
                        &c->gx32};
  for (auto *b : fbs) b->release();
  DBuf<int> *ibs[] = {&c->tpa, &c->tpb, &c->cell_nodes, &c->row_ptr, &c->col, &c->diag, &c->cell_blk,
                      &c->color_cells, &c->status};
  for (auto *b : ibs) b->release();
  c->cell_face.release();
  c->node_mask.release();
  for (auto e : c->ev_pool)
    if (e) (void)hipEventDestroy(e);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return GLH_OK;
}

const char *glh_last_error(const glh_ctx *c) { return c ? c->err.c_str() : "null context"; }

int glh_set_material(glh_ctx *c, double alpha, const double beta[5], const double K[3], double bt)
{
  if (!c || !beta || !K) return GLH_ERR_ARG;
  if (!(bt > 0)) return fail(c, GLH_ERR_ARG, "AdGR diffuse length must be positive");
  c->mat.al = alpha;
  c->mat.b1 = beta[0]; c->mat.b2 = beta[1]; c->mat.b3 = beta[2]; c->mat.b4 = beta[3]; c->mat.b5 = beta[4];
  c->mat.K1 = K[0];
  c->mat.K23 = K[1] + K[2];
  c->mat.robin = (bt < 1e10) ? K[0] / bt : 0.0;
  c->have_mat = true;
  c->j_valid = c->d_valid = false;
  return GLH_OK;
}

int glh_mesh_box(int degree, int nx, int ny, int nz, const double lo[3], const double hi[3], const int side_id[6],
                 glh_mesh **out)
{
  if (!out || !lo || !hi || !side_id) return GLH_ERR_ARG;
  *out = nullptr;
  glh_mesh *m = new glh_mesh();
  std::string err;
  if (glh::make_box(m->m, degree, nx, ny, nz, lo, hi, side_id, err)) {
    delete m;
    return GLH_ERR_ARG;
  }
  *out = m;
  return GLH_OK;
}

int glh_mesh_destroy(glh_mesh *m)
{
  delete m;
  return GLH_OK;
}

int glh_mesh_sizes(const glh_mesh *mm, int64_t sizes[8])
{
  if (!mm || !sizes) return GLH_ERR_ARG;
  const glh::Mesh &m = mm->m;
  sizes[0] = m.deg; sizes[1] = m.n_cells; sizes[2] = m.n_nodes; sizes[3] = (int64_t)m.faces.size() / 3;
  sizes[4] = m.row_ptr.empty() ? 0 : m.row_ptr.back(); sizes[5] = (int64_t)m.color_ptr.size() - 1;
  sizes[6] = 0; sizes[7] = 0;
  return GLH_OK;
}

int glh_mesh_get(const glh_mesh *mm, int *cell_nodes, double *cell_h, double *node_xyz, int *faces,
                 uint32_t *node_mask, int *color_ptr, int *color_cells, int *row_ptr, int *col)
{
  if (!mm) return GLH_ERR_ARG;
  const glh::Mesh &m = mm->m;
  auto cp = [](auto *dst, const auto &v) {
    if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
  };
  cp(cell_nodes, m.cell_nodes); cp(cell_h, m.cell_h); cp(node_xyz, m.node_xyz); cp(faces, m.faces);
  cp(node_mask, m.node_mask); cp(color_ptr, m.color_ptr); cp(color_cells, m.color_cells); cp(row_ptr, m.row_ptr);
  cp(col, m.col);
  return GLH_OK;
}

int glh_upload_mesh(glh_ctx *c, const glh_mesh *mm)
{
  if (!c || !mm) return GLH_ERR_ARG;
  const glh::Mesh &m = mm->m;
  HIPCHK(c, hipSetDevice(c->device));
  if (c->deg != m.deg || !c->tN.p) {
    int rc = build_tables(c, m.deg);
    if (rc) return rc;
  }
  hipStream_t s = c->stream;
  c->n_cells = m.n_cells;
  c->n_nodes = m.n_nodes;
  c->nnzb = m.row_ptr.back();
  c->color_ptr = m.color_ptr;
  // per-cell face codes (boundary id + 1, 3 bits per local face)
  std::vector<uint32_t> cface(m.n_cells, 0u);
  for (size_t f = 0; f + 2 < m.faces.size(); f += 3) {
    const int e = m.faces[f], lf = m.faces[f + 1], bid = m.faces[f + 2];
    const uint32_t code = (uint32_t)std::min(bid, 6) + 1u;
    cface[e] |= code << (3 * lf);
  }
  std::vector<int> diag(m.n_nodes);
  for (int r = 0; r < m.n_nodes; ++r) {
    diag[r] = glh::find_block(m, r, r);
    if (diag[r] < 0) return fail(c, GLH_ERR_ARG, "pattern lacks a diagonal block");
  }
  // cell -> block map with first-writer marks in colour order
  const int ns = m.ns;
  std::vector<int> cblk((size_t)m.n_cells * ns * ns);
  std::vector<uint8_t> touched(c->nnzb, 0);
  for (size_t k = 0; k + 1 < m.color_ptr.size(); ++k)
    for (int i = m.color_ptr[k]; i < m.color_ptr[k + 1]; ++i) {
      const int e = m.color_cells[i];
      const int *cn = &m.cell_nodes[(size_t)e * ns];
      for (int a = 0; a < ns; ++a)
        for (int b = 0; b < ns; ++b) {
          const int blk = glh::find_block(m, cn[a], cn[b]);
          if (blk < 0) return fail(c, GLH_ERR_ARG, "pattern lacks a cell coupling");
          const bool first = !touched[blk];
          touched[blk] = 1;
          cblk[((size_t)e * ns + a) * ns + b] = first ? -(blk + 1) : blk;
        }
    }
  HIPCHK(c, upload(c->cell_nodes, m.cell_nodes.data(), m.cell_nodes.size(), s));
  HIPCHK(c, upload(c->cell_h, m.cell_h.data(), m.cell_h.size(), s));
  HIPCHK(c, upload(c->cell_face, cface.data(), cface.size(), s));
  HIPCHK(c, upload(c->node_mask, m.node_mask.data(), m.node_mask.size(), s));
  HIPCHK(c, upload(c->row_ptr, m.row_ptr.data(), m.row_ptr.size(), s));
  HIPCHK(c, upload(c->col, m.col.data(), m.col.size(), s));
  HIPCHK(c, upload(c->diag, diag.data(), diag.size(), s));
  HIPCHK(c, upload(c->cell_blk, cblk.data(), cblk.size(), s));
  HIPCHK(c, upload(c->color_cells, m.color_cells.data(), m.color_cells.size(), s));
  const size_t n = (size_t)m.n_nodes * NC;
  HIPCHK(c, c->U.alloc(n));
  HIPCHK(c, c->dU.alloc(n));
  HIPCHK(c, c->X.alloc(n));
  HIPCHK(c, c->rhs.alloc(n));
  HIPCHK(c, c->res.alloc(n));
  HIPCHK(c, c->Fcell.alloc(m.n_cells));
  HIPCHK(c, c->red_partial.alloc(glh::MDOT_GRID * 2));
  HIPCHK(c, c->red_out.alloc(4));
  HIPCHK(c, hipMemsetAsync(c->U.p, 0, n * sizeof(double), s));
  HIPCHK(c, hipMemsetAsync(c->dU.p, 0, n * sizeof(double), s));
  HIPCHK(c, hipMemsetAsync(c->rhs.p, 0, n * sizeof(double), s));
  HIPCHK(c, hipStreamSynchronize(s));
  c->j_valid = c->d_valid = false;
  c->have_mesh = true;
  return GLH_OK;
}

int glh_set_solution(glh_ctx *c, const double *Uh)
{
  int rc = check_ready(c, false);
  if (rc) return rc;
  if (!Uh) return fail(c, GLH_ERR_ARG, "null solution");
  HIPCHK(c, hipMemcpyAsync(c->U.p, Uh, ndofs(c) * sizeof(double), hipMemcpyHostToDevice, c->stream));
  // constraints_solution.distribute (setup_uniform_B-phase.cc:262)
  HIPCHK(c, glh::launch_mask_zero(c->n_nodes, c->node_mask.p, c->U.p, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return GLH_OK;
}

int glh_store_initial(glh_ctx *c)
{
  int rc = check_ready(c, false);
  if (rc) return rc;
  HIPCHK(c, c->U0.alloc(ndofs(c)));
  HIPCHK(c, hipMemcpyAsync(c->U0.p, c->U.p, ndofs(c) * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return GLH_OK;
}

int glh_restore_initial(glh_ctx *c)
{
  int rc = check_ready(c, false);
  if (rc) return rc;
  if (!c->U0.p || c->U0.n < ndofs(c)) return fail(c, GLH_ERR_STATE, "no stored initial state");
  HIPCHK(c, hipMemcpyAsync(c->U.p, c->U0.p, ndofs(c) * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return GLH_OK;
}

static int get_vec(glh_ctx *c, const double *src, double *dst)
{
  int rc = check_ready(c, false);
  if (rc) return rc;
  if (!dst) return fail(c, GLH_ERR_ARG, "null output");
  HIPCHK(c, hipMemcpyAsync(dst, src, ndofs(c) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return GLH_OK;
}

int glh_get_solution(glh_ctx *c, double *Uh) { return c ? get_vec(c, c->U.p, Uh) : GLH_ERR_ARG; }
int glh_get_update(glh_ctx *c, double *d) { return c ? get_vec(c, c->dU.p, d) : GLH_ERR_ARG; }
int glh_get_rhs(glh_ctx *c, double *r) { return c ? get_vec(c, c->rhs.p, r) : GLH_ERR_ARG; }

int glh_get_jacobian(glh_ctx *c, double *val)
{
  int rc = check_ready(c, false);
  if (rc) return rc;
  if (!c->j_valid) return fail(c, GLH_ERR_STATE, "no Jacobian assembled");
  const size_t nv = (size_t)c->nnzb * 324;
  if (!c->j_fp32) {
    HIPCHK(c, hipMemcpyAsync(val, c->J64.p, nv * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  } else {
    std::vector<float> tmp(nv);
    HIPCHK(c, hipMemcpyAsync(tmp.data(), c->J32.p, nv * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < nv; ++i) val[i] = tmp[i];
  }
  return GLH_OK;
}

int glh_assemble(glh_ctx *c, int flags)
{
  int rc = check_ready(c);
  if (rc) return rc;
  const bool dj = flags & GLH_ASM_J, dr = flags & GLH_ASM_R, f32 = flags & GLH_ASM_FP32;
  if (!dj && !dr) return fail(c, GLH_ERR_ARG, "nothing to assemble");
  rc = assemble_at(c, c->U.p, dj, dr, f32, c->rhs.p, true);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return GLH_OK;
}

int glh_rhs_norm(glh_ctx *c, double *nrm)
{
  int rc = check_ready(c);
  if (rc) return rc;
  return norm64(c, c->rhs.p, nrm);
}

int glh_residual(glh_ctx *c, double *nrm)
{
  int rc = check_ready(c);
  if (rc) return rc;
  rc = assemble_at(c, c->U.p, false, true, false, c->res.p, false);
  if (rc) return rc;
  return norm64(c, c->res.p, nrm);
}

int glh_solve(glh_ctx *c, double rtol, int maxit, int restart, double stats[2])
{
  int rc = check_ready(c);
  if (rc) return rc;
  return solve_impl(c, rtol, maxit, restart, stats);
}

int glh_newton_step(glh_ctx *c, double step, int damped, double out[4])
{
  int rc = check_ready(c);
  if (rc) return rc;
  if (damped && !(step > 0 && step < 1)) return fail(c, GLH_ERR_ARG, "line-search step must be in (0,1)");
  return newton_step_impl(c, step, damped, out);
}

int glh_free_energy(glh_ctx *c, double *F)
{
  int rc = check_ready(c);
  if (rc) return rc;
  if (!F) return fail(c, GLH_ERR_ARG, "null output");
  HIPCHK(c, glh::launch_energy(c->deg, c->tables(), c->meshdev(), c->mat, c->U.p, c->n_cells, c->Fcell.p,
                               c->stream));
  HIPCHK(c, c->red_partial.alloc(glh::MDOT_GRID * 2));
  HIPCHK(c, glh::launch_sum(c->n_cells, c->Fcell.p, c->red_partial.p, c->red_out.p, c->stream));
  HIPCHK(c, hipMemcpyAsync(F, c->red_out.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return GLH_OK;
}

int glh_newton_solve(glh_ctx *c, double rtol, int max_newton, double lin_tol, int lin_maxit, int restart,
                     double ls_step, int damped, int fp32, double stats[10])
{
  int rc = check_ready(c);
  if (rc) return rc;
  using clk = std::chrono::steady_clock;
  double t_asm = 0, t_sol = 0, t_ls = 0;
  int lin_total = 0, k = 0, trials = 0;
  auto t0 = clk::now();
  rc = assemble_at(c, c->U.p, false, true, false, c->rhs.p, false);
  if (rc) return rc;
  double r0 = 0;
  rc = norm64(c, c->rhs.p, &r0);
  if (rc) return rc;
  t_asm += std::chrono::duration<double>(clk::now() - t0).count();
  double rcur = r0;
  const bool f32 = fp32 != 0;
  for (k = 0; k < max_newton; ++k) {
    if (rcur <= rtol * r0) break;
    t0 = clk::now();
    rc = assemble_at(c, c->U.p, true, true, f32, c->rhs.p, false);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    auto t1 = clk::now();
    t_asm += std::chrono::duration<double>(t1 - t0).count();
    double st[2] = {0, 0};
    rc = solve_impl(c, lin_tol, lin_maxit, restart, st);
    lin_total += (int)st[0];
    if (rc) return rc;
    auto t2 = clk::now();
    t_sol += std::chrono::duration<double>(t2 - t1).count();
    double out[4];
    rc = newton_step_impl(c, ls_step, damped, out);
    if (rc) return rc;
    rcur = out[1];
    trials += (int)out[3];
    t_ls += std::chrono::duration<double>(clk::now() - t2).count();
  }
  double F = 0;
  rc = glh_free_energy(c, &F);
  if (rc) return rc;
  if (stats) {
    stats[0] = k; stats[1] = lin_total; stats[2] = r0; stats[3] = rcur;
    stats[4] = t_asm; stats[5] = t_sol; stats[6] = t_ls; stats[7] = F;
    stats[8] = trials; stats[9] = (rcur <= rtol * r0) ? 1.0 : 0.0;
  }
  return GLH_OK;
}

int glh_last_assembly_ms(glh_ctx *c, double ms[2])
{
  if (!c || !ms) return GLH_ERR_ARG;
  ms[0] = c->asm_ms;
  ms[1] = c->asm_launches;
  return GLH_OK;
}

int glh_set_timing(glh_ctx *c, int on)
{
  if (!c) return GLH_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  if (on && c->ev_pool.empty()) {
    c->ev_pool.resize(2 * 1024, nullptr);
    for (auto &e : c->ev_pool) HIPCHK(c, hipEventCreate(&e));
  }
  c->timing = on != 0;
  c->ev_used = 0;
  c->spmv_ms_total = c->spmv_count = 0;
  c->jasm_ms_total = c->jasm_count = c->rasm_ms_total = c->rasm_count = 0;
  return GLH_OK;
}

int glh_timing(glh_ctx *c, double out[6])
{
  if (!c || !out) return GLH_ERR_ARG;
  int rc = flush_events(c);
  if (rc) return rc;
  out[0] = c->spmv_ms_total; out[1] = c->spmv_count;
  out[2] = c->jasm_ms_total; out[3] = c->jasm_count;
  out[4] = c->rasm_ms_total; out[5] = c->rasm_count;
  return GLH_OK;
}

int glh_bench_spmv(glh_ctx *c, int k, double *ms_avg)
{
  int rc = check_ready(c);
  if (rc) return rc;
  if (!c->j_valid || k < 1 || !ms_avg) return fail(c, GLH_ERR_STATE, "no Jacobian assembled");
  HIPCHK(c, hipEventRecord(c->ev0, c->stream));
  for (int i = 0; i < k; ++i) {
    if (c->j_fp32) {
      HIPCHK(c, c->gW32.alloc(ndofs(c)));
      HIPCHK(c, c->gZ32.alloc(ndofs(c)));
      if (i == 0) HIPCHK(c, glh::launch_convert<double, float>(ndofs(c), c->U.p, c->gZ32.p, c->stream));
      HIPCHK(c, glh::launch_spmv<float, float>(c->n_nodes, c->row_ptr.p, c->col.p, c->J32.p, c->gZ32.p, c->gW32.p,
                                               c->stream));
    } else {
      HIPCHK(c, glh::launch_spmv<double, double>(c->n_nodes, c->row_ptr.p, c->col.p, c->J64.p, c->U.p, c->res.p,
                                                 c->stream));
    }
  }
  HIPCHK(c, hipEventRecord(c->ev1, c->stream));
  HIPCHK(c, hipEventSynchronize(c->ev1));
  float ms = 0;
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
  *ms_avg = ms / k;
  return GLH_OK;
}

int glh_bench_copy(glh_ctx *c, size_t nbytes, int k, double *ms_avg)
{
  if (!c || !ms_avg || k < 1 || nbytes % 16) return GLH_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->device));
  void *a = nullptr, *b = nullptr;
  HIPCHK(c, hipMalloc(&a, nbytes));
  hipError_t e = hipMalloc(&b, nbytes);
  if (e != hipSuccess) {
    (void)hipFree(a);
    return fail(c, GLH_ERR_HIP, "hipMalloc failed");
  }
  HIPCHK(c, hipMemsetAsync(a, 0, nbytes, c->stream));
  HIPCHK(c, hipEventRecord(c->ev0, c->stream));
  for (int i = 0; i < k; ++i) HIPCHK(c, glh::launch_copy(nbytes, a, b, c->stream));
  HIPCHK(c, hipEventRecord(c->ev1, c->stream));
  HIPCHK(c, hipEventSynchronize(c->ev1));
  float ms = 0;
  HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
  *ms_avg = ms / k;
  (void)hipFree(a);
  (void)hipFree(b);
  return GLH_OK;
}

}  // extern "C"
