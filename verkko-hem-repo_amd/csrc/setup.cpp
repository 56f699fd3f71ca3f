// Seeded initial guesses (the setup_system() initial data of
// femgl/src/setup_uniform_B-phase.cc:200-265 and setup_uniform_A-phase.cc:
// 225-262).  The reference draws from std::random_device, which cannot be
// reproduced; here the bulk phase is perturbed by
// amp * gap * (2 U - 1), U = (mt19937_64() >> 11) * 2^-53, drawn for every
// dof in node-major order, and the Dirichlet components are then zeroed
// (constraints_solution.distribute, setup_uniform_B-phase.cc:262).
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
