// The adaptive cycle loop of FemGL::run (femgl/src/run.cc:109-260) over the
// C ABI: for every cycle, the mesh of the current forest is uploaded, the
// Newton solve runs on the GPU, then the Kelly indicators of the converged
// field flag the fixed fraction of cells (refine.cc:146-157), the forest is
// refined with 2:1 balance and the field is interpolated onto the new mesh
// on the device (refine.cc:163-170).  Cycle 0 starts from the caller's guess.
//
// The reference stops each cycle on an absolute ||R|| or a stagnation test
// (run.cc:236-255); here every cycle converges to ||R||/||R0|| <= rtol, the
// criterion of the benchmark workloads.  With check_fp64 a mixed-precision
// run also solves every cycle in FP64 from the same start and reports the
// difference of the two fixed points.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <vector>

#include "glhip.h"
#include "handles.hpp"

namespace {

double seconds()
{
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

double max_abs(const std::vector<double> &v)
{
  double m = 0.0;
  for (double x : v) m = std::max(m, std::fabs(x));
  return m;
}

}  // namespace

extern "C" int glh_run_adaptive(glh_ctx *ctx, const glh_forest *f0, const double *U0, int n_refine,
                                double refine_frac, double rtol, int max_newton, double lin_tol, int lin_maxit,
                                int restart, double ls_step, int damped, int fp32, int check_fp64,
                                double *cycle_stats, glh_forest **f_out, glh_mesh **m_out)
{
  if (!ctx || !f0 || !U0 || n_refine < 0 || !(refine_frac >= 0 && refine_frac <= 1) || !cycle_stats)
    return GLH_ERR_ARG;
  if (f_out) *f_out = nullptr;
  if (m_out) *m_out = nullptr;
  const double nan = std::numeric_limits<double>::quiet_NaN();
  glh_forest *f = new glh_forest(*f0);
  glh_mesh *m = nullptr;
  // mesh of the current cycle; from cycle 1 on it is uploaded, with the
  // field transferred on the device, at the end of the previous cycle
  glh_mesh *next = nullptr;
  int rc = GLH_OK;
  for (int cyc = 0; cyc <= n_refine; ++cyc) {
    const double t0 = seconds();
    double *st = cycle_stats + (size_t)cyc * GLH_CYCLE_STATS;
    for (int k = 0; k < GLH_CYCLE_STATS; ++k) st[k] = nan;
    glh_mesh *mn = next;
    next = nullptr;
    if (cyc == 0) {
      if ((rc = glh_forest_mesh(f, &mn)) != GLH_OK) break;
      if ((rc = glh_upload_mesh(ctx, mn)) != GLH_OK || (rc = glh_set_solution(ctx, U0)) != GLH_OK) {
        glh_mesh_destroy(mn);
        break;
      }
    }
    const glh::Mesh &M = mn->m;
    if ((rc = glh_store_initial(ctx)) != GLH_OK) {
      glh_mesh_destroy(mn);
      break;
    }
    double s[10];
    const double ts = seconds();
    if ((rc = glh_newton_solve(ctx, rtol, max_newton, lin_tol, lin_maxit, restart, ls_step, damped, fp32, s)) !=
        GLH_OK) {
      glh_mesh_destroy(mn);
      break;
    }
    const double t_solve = seconds() - ts;
    std::vector<double> Uc((size_t)M.n_nodes * 18);
    if ((rc = glh_get_solution(ctx, Uc.data())) != GLH_OK) {
      glh_mesh_destroy(mn);
      break;
    }
    int64_t nh = 0;
    for (int v = 0; M.has_hanging() && v < M.n_nodes; ++v) nh += M.hang_ptr[v + 1] > M.hang_ptr[v];
    st[0] = cyc; st[1] = M.n_cells; st[2] = (double)M.n_nodes * 18; st[3] = (double)nh;
    st[4] = s[0]; st[5] = s[1]; st[6] = s[2]; st[7] = s[3]; st[8] = s[7]; st[9] = s[9]; st[10] = t_solve;
    if (check_fp64 && fp32) {
      double s64[10];
      std::vector<double> U64(Uc.size());
      if ((rc = glh_restore_initial(ctx)) != GLH_OK ||
          (rc = glh_newton_solve(ctx, rtol, max_newton, lin_tol, lin_maxit, restart, ls_step, damped, 0, s64)) !=
              GLH_OK ||
          (rc = glh_get_solution(ctx, U64.data())) != GLH_OK) {
        glh_mesh_destroy(mn);
        break;
      }
      double du = 0.0;
      for (size_t i = 0; i < U64.size(); ++i) du = std::max(du, std::fabs(Uc[i] - U64[i]));
      st[12] = s64[7];
      st[13] = std::fabs(s[7] - s64[7]) / std::max(std::fabs(s64[7]), 1e-300);
      st[14] = du / std::max(max_abs(U64), 1e-300);
      st[15] = s64[9];
      // continue from the mixed-precision field, as the run without the check
      if ((rc = glh_set_solution(ctx, Uc.data())) != GLH_OK) {
        glh_mesh_destroy(mn);
        break;
      }
    }
    if (cyc < n_refine) {
      std::vector<float> eta(M.n_cells);
      std::vector<uint8_t> fl(M.n_cells);
      glh_forest *fn = nullptr;
      glh_mesh *mnext = nullptr;
      if ((rc = glh_kelly(f, mn, Uc.data(), eta.data())) != GLH_OK ||
          (rc = glh_flag_fixed_number(M.n_cells, eta.data(), refine_frac, fl.data())) != GLH_OK ||
          (rc = glh_forest_refine(f, fl.data(), &fn)) != GLH_OK || (rc = glh_forest_mesh(fn, &mnext)) != GLH_OK) {
        if (fn) glh_forest_destroy(fn);
        glh_mesh_destroy(mn);
        break;
      }
      // the context holds Uc: refine.cc:163-175 on the device
      rc = glh_transfer_solution(ctx, f, mn, fn, mnext);
      glh_mesh_destroy(mn);
      glh_forest_destroy(f);
      f = fn;
      if (rc != GLH_OK) {
        glh_mesh_destroy(mnext);
        break;
      }
      next = mnext;
    } else {
      m = mn;
    }
    st[11] = seconds() - t0;
  }
  if (rc != GLH_OK) {
    glh_forest_destroy(f);
    if (m) glh_mesh_destroy(m);
    return rc;
  }
  if (f_out) *f_out = f;
  else glh_forest_destroy(f);
  if (m_out) *m_out = m;
  else if (m) glh_mesh_destroy(m);
  return GLH_OK;
}
