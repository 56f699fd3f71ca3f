// Seeded initial guesses (the setup_system() initial data of
// femgl/src/setup_uniform_B-phase.cc:200-265 and setup_uniform_A-phase.cc:
// 225-262).  The reference draws from std::random_device, which cannot be
// reproduced; here the bulk phase is perturbed by
// amp * gap * (2 U - 1), U = (mt19937_64() >> 11) * 2^-53, drawn for every
// dof in node-major order, and the Dirichlet components are then zeroed
// (constraints_solution.distribute, setup_uniform_B-phase.cc:262).
#include <cmath>
#include <cstdint>
#include <random>
#include <vector>

#include "glhip.h"

extern "C" int glh_initial_guess_host(int n_nodes, double gap, int phase, double amp, uint64_t seed, double *U)
{
  if (!U || n_nodes < 0 || (phase != 0 && phase != 1)) return GLH_ERR_ARG;
  std::mt19937_64 gen(seed);
  // bulk B phase: u_11 = u_22 = u_33 = gap * 0.577350269f (setup_uniform_B-phase.cc:245-259)
  // bulk A phase: u_11 = v_12 = gap * 0.707107f (setup_uniform_A-phase.cc:238-250)
  const double fB = (double)0.577350269f, fA = (double)0.707107f;
  for (int64_t nd = 0; nd < n_nodes; ++nd)
    for (int c = 0; c < 18; ++c) {
      double base = 0.0;
      if (phase == 0 && (c == 0 || c == 4 || c == 8)) base = gap * fB;
      if (phase == 1 && (c == 0 || c == 10)) base = gap * fA;
      const double r = (double)(gen() >> 11) * (1.0 / 9007199254740992.0);
      U[nd * 18 + c] = base + amp * gap * (2.0 * r - 1.0);
    }
  return GLH_OK;
}

// A-phase / B-phase flat-wall guess (BnA of femgl/inc/BnA.h:130-163 with
// femgl/src/setup_uniform_BnA-flatwall-configuration.cc:228-235): z >= z_a
// takes the bulk B phase (u_11 = u_22 = u_33 = gap_B * 0.5773502691896258f),
// z < z_a the bulk A phase (u_11 = v_12 = gap_A * 0.7071067811865475f); then
// amp * gap_B * (2 U - 1) noise from the stream above.
extern "C" int glh_initial_guess_bna(int n_nodes, const double *node_xyz, double z_a, double gap_A, double gap_B,
                                     double amp, uint64_t seed, double *U)
{
  if (!U || !node_xyz || n_nodes < 0) return GLH_ERR_ARG;
  std::mt19937_64 gen(seed);
  const double fB = (double)0.5773502691896258f, fA = (double)0.7071067811865475f;
  for (int64_t nd = 0; nd < n_nodes; ++nd) {
    const bool b = node_xyz[3 * nd + 2] >= z_a;
    for (int c = 0; c < 18; ++c) {
      double base = 0.0;
      if (b && (c == 0 || c == 4 || c == 8)) base = gap_B * fB;
      if (!b && (c == 0 || c == 10)) base = gap_A * fA;
      const double r = (double)(gen() >> 11) * (1.0 / 9007199254740992.0);
      U[nd * 18 + c] = base + amp * gap_B * (2.0 * r - 1.0);
    }
  }
  return GLH_OK;
}

namespace {

// z-profiles of the weak-coupling PDW ansatz (femgl/inc/pdw_config.h:183-299),
// in the reduced coordinate s = (z + half_D) / D, evaluated by Horner's rule.
double horner(const double *c, int n, double s)
{
  double r = c[n - 1];
  for (int k = n - 2; k >= 0; --k) r = r * s + c[k];
  return r;
}
const double PDW_P6[7] = {-0.010674533968153196, -0.04928146307995591, 0.4123298769345518, -5.957210687551436,
                          15.697110273856326,    -14.963347649994844, 4.860888318571941};
const double PDW_P9[10] = {0.005381764646081998, -0.09659207269737603, 3.059717163195402,  -35.01132629189211,
                           193.9107091164305,    -594.3817533999686,   1058.3851407510472, -1088.4440880311813,
                           599.3158862587954,    -136.74038236803835};
const double PDW_P7[8] = {-0.008476309862633497, -0.15843411116780842, 1.1880072170146772, -8.79341508668322,
                          32.93035389045762,     -56.473628486426286,  44.90526146438388,  -13.60443812339298};
const double PDW_P3[4] = {0.703221673991666, 0.07213025175696147, -4.458423083418102, 2.9780246543754445};

// B-n-B polka-dot guess (femgl/inc/BinB.h:130-162): fixed u = v matrices
// inside and outside the cylinder x^2 + y^2 <= r^2.
const double BNB_IN[9] = {0.859734, -0.285668, -1.76655, 1.58845, -1.25931, 0.763255, 1.26923, 1.77689, 0.245094};
const double BNB_OUT[9] = {0.668541, 0.632763, 1.7257, 1.63496, -1.50349, -0.0439564, 1.3531, 1.50482, -0.799212};

}  // namespace

// Analytic initial configurations interpolated at the nodes (deal.II's
// VectorTools::interpolate evaluates a Function at the FE_Q support points,
// which are the mesh nodes), plus amp * scale * (2 U - 1) noise from the
// stream of glh_initial_guess_host (amp = 0 reproduces the reference, which
// adds none).  par by kind:
//   GLH_GUESS_BINA (femgl/inc/BinA.h:130-161, setup_uniform_BinA-bubble-like-
//     configuration.cc:229-235): {radius, gap_A, gap_B}; B phase inside the
//     ball |x| <= radius, A phase outside; scale gap_B.
//   GLH_GUESS_BNB (BinB.h, setup_uniform_BnB-polkka-dot-like-configuration.cc:
//     309-313): {radius}; scale 1.
//   GLH_GUESS_PDW (pdw_config.h:124-299, setup_weak-coupling-PDW-
//     configuration.cc:311-315): {u, half_D, gap_B}; Q = pi / (sqrt(3) 2 half_D).
extern "C" int glh_initial_guess_field(int kind, int n_nodes, const double *node_xyz, const double *par, double amp,
                                       uint64_t seed, double *U)
{
  if (!U || !node_xyz || !par || n_nodes < 0) return GLH_ERR_ARG;
  if (kind != GLH_GUESS_BINA && kind != GLH_GUESS_BNB && kind != GLH_GUESS_PDW) return GLH_ERR_ARG;
  std::mt19937_64 gen(seed);
  const double s3 = std::sqrt(3.0), s2 = std::sqrt(2.0), s6 = std::sqrt(6.0);
  double scale = 1.0;
  if (kind == GLH_GUESS_BINA) scale = par[2];
  if (kind == GLH_GUESS_PDW) scale = par[2];
  for (int64_t nd = 0; nd < n_nodes; ++nd) {
    const double x = node_xyz[3 * nd], y = node_xyz[3 * nd + 1], z = node_xyz[3 * nd + 2];
    double v[18] = {0};
    if (kind == GLH_GUESS_BINA) {
      const double r2 = par[0] * par[0];
      if (x * x + y * y + z * z <= r2) {
        v[0] = v[4] = v[8] = par[2] * (double)0.5773502691896258f;
      } else {
        v[0] = v[10] = par[1] * (double)0.7071067811865475f;
      }
    } else if (kind == GLH_GUESS_BNB) {
      const double *t = (x * x + y * y <= par[0] * par[0]) ? BNB_IN : BNB_OUT;
      for (int c = 0; c < 9; ++c) v[c] = v[9 + c] = t[c];
    } else {
      const double u = par[0], hD = par[1], gap = par[2], D = 2.0 * hD;
      const double q = 3.141592653589793 / (1.7320508075688772 * 2.0 * hD);
      const double s = (z + hD) / D;
      const double p6 = horner(PDW_P6, 7, s), p9 = horner(PDW_P9, 10, s), p7 = horner(PDW_P7, 8, s),
                   p3 = horner(PDW_P3, 4, s);
      const double a = q * x, b = s3 * q * y / 2.0;
      const double ca = std::cos(a), ch = std::cos(a / 2.0), sh = std::sin(a / 2.0), cb = std::cos(b),
                   sb = std::sin(b), sa = std::sin(a);
      const double c1 = ca - ch * cb, c2 = ca + 2.0 * ch * cb, sA = sh * sb, sB = cb * sh + sa, sC = ch * sb;
      const double g = gap * u, b0 = gap / s3;
      v[0] = b0 + g * (2.0 * p6 + 2.0 * p9 * c1);
      v[1] = g * (2.0 * p6 - 2.0 * s3 * p9 * sA);
      v[2] = g * (2.0 * p6 + 2.0 * s2 * p7 * sB);
      v[3] = g * (2.0 * s3 * p6 * c2 - 2.0 * s3 * p9 * sA);
      v[4] = b0 + g * (-2.0 * p9 * c1 + 2.0 * s3 * p6 * c2);
      v[5] = g * (2.0 * s3 * p6 * c2 + 2.0 * s6 * p7 * sC);
      v[6] = g * (2.0 * s2 * p3 * sB);
      v[7] = g * (2.0 * s6 * p3 * sC);
      v[8] = b0;
    }
    for (int c = 0; c < 18; ++c) {
      const double r = (double)(gen() >> 11) * (1.0 / 9007199254740992.0);
      U[nd * 18 + c] = v[c] + amp * scale * (2.0 * r - 1.0);
    }
  }
  return GLH_OK;
}
