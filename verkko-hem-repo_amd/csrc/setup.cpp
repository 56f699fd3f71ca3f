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
