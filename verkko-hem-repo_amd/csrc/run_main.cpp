// glhip_run: the solver executable (the role of sol/src/main.cc:101-148).
// Reads a parameter file (default ./configuration.prm, as main.cc:112),
// runs the cycle loop of FemGL::run on GPU 0 and prints one line per
// refinement cycle.  Errors end the program with status 1 and a message on
// stderr, as the reference's exception handlers do (main.cc:120-145).
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "glhip.h"

static int abort_with(const std::string &what)
{
  std::fprintf(stderr, "\n\n----------------------------------------------------\n"
                       "Exception on processing: \n%s\nAborting!\n"
                       "----------------------------------------------------\n",
               what.c_str());
  return 1;
}

int main(int argc, char **argv)
{
  const std::string path = argc > 1 ? argv[1] : "configuration.prm";
  if (path == "-h" || path == "--help") {
    std::printf("usage: glhip_run [configuration.prm]\n");
    return 0;
  }
  std::ifstream in(path);
  if (!in) return abort_with("cannot open parameter file " + path);
  std::stringstream ss;
  ss << in.rdbuf();
  char err[1024] = {0};
  glh_prm *prm = nullptr;
  if (glh_prm_parse(ss.str().c_str(), &prm, err, (int)sizeof(err)) != GLH_OK)
    return abort_with(std::string("parameter file ") + path + ": " + err);
  glh_ctx *ctx = nullptr;
  if (glh_create(0, &ctx) != GLH_OK) {
    glh_prm_destroy(prm);
    return abort_with("no usable GPU (glh_create failed)");
  }
  std::vector<double> rows(64 * 6);
  int n = 0;
  const int rc = glh_run_prm(ctx, prm, 64, rows.data(), &n);
  std::string msg = rc ? std::string(glh_last_error(ctx)) : std::string();
  for (int k = 0; k < n; ++k) {
    const double *r = &rows[(size_t)k * 6];
    std::printf("cycle %d: %d cells, %d dofs, %d Newton iterations, ||R|| %.6e, free energy %.12g\n", (int)r[0],
                (int)r[1], (int)r[2], (int)r[3], r[4], r[5]);
  }
  glh_destroy(ctx);
  glh_prm_destroy(prm);
  if (rc) return abort_with("glh_run_prm failed (" + std::to_string(rc) + ")" + (msg.empty() ? "" : ": " + msg));
  return 0;
}
