// The reference's .prm configuration (deal.II ParameterHandler text: nested
// `subsection NAME` ... `end` blocks with `set KEY = VALUE` lines, `#`
// comments) with the keys, defaults and patterns of
// confreader/src/declare.cc:110-300, plus the build's own subsection
// "glhip" for what the reference fixes at compile time (geometry variant,
// boundary ids, element degree, initial guess, seed, precision).
//
// glh_run_prm restates FemGL::run (femgl/src/run.cc:109-260) on top of the
// C ABI: make_grid (hyper_rectangle or hyper_cube refined globally), per
// cycle up to n_iteration+1 rounds of assemble / solve(cycle tol) / damped
// Newton step, the stuck-refine test |R - R_last| < threshold, the final
// absolute test ||R|| <= converge accuracy, and adaptive (Kelly, fixed
// fraction) or global refinement between cycles.
#include <sys/stat.h>

#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "glhip.h"
#include "handles.hpp"

namespace {

struct Entry {
  std::string def;
  char type;       // 'd' double, 'i' integer, 'b' bool, 's' string
  bool nonneg;     // Patterns::Double(0) / Integer(0)
  std::string value;
};

std::string trim(const std::string &s)
{
  size_t a = 0, b = s.size();
  while (a < b && std::isspace((unsigned char)s[a])) ++a;
  while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
  return s.substr(a, b - a);
}

void declare(std::map<std::string, Entry> &m)
{
  auto d = [&](const std::string &sub, const std::string &key, const std::string &def, char type,
               bool nonneg = true) { m[sub + "/" + key] = Entry{def, type, nonneg, def}; };
  const std::string P = "physical parameters", C = "control parameters", X = "glhip";
  // confreader/src/declare.cc:115-149
  d(P, "pressure in bar", "0.0", 'd');
  d(P, "t_reduced", "0.0", 'd');
  d(P, "AdGR diffuse length", "1.0e10", 'd');
  d(P, "gaussian random mean value", "2.0", 'd');
  d(P, "gaussian random STD", "0.1", 'd');
  d(P, "trun on Strong Coupling Correction", "true", 'b');
  d(P, "amplitude u parameter", "1.0", 'd', false);
  // confreader/src/declare.cc:155-291
  d(C, "cube half side length", "20", 'd');
  d(C, "half x length of retangle", "20", 'd');
  d(C, "half y length of retangle", "20", 'd');
  d(C, "half z length of retangle", "20", 'd');
  d(C, "B-phase inner plate radius ratio", "0.4666", 'd');
  d(C, "B-phase ball radius ratio", "0.5", 'd');
  d(C, "A-phase block range ratio", "0.0", 'd');
  d(C, "Number of refinements", "4", 'i');
  d(C, "Number of interations", "40", 'i');
  for (int k = 0; k < 4; ++k) d(C, "Cycle " + std::to_string(k) + " refinement threshold", "1.0e0", 'd');
  for (int k = 1; k < 5; ++k) d(C, "Cycle " + std::to_string(k) + " do global refinement", "false", 'b');
  for (int k = 0; k < 5; ++k) d(C, "Cycle " + std::to_string(k) + " linear solver tol", "1.0e-1", 'd');
  d(C, "converge accuracy", "5.0e-6", 'd');
  d(C, "adaptive refinment ratio", "0.3", 'd');
  d(C, "adaptive coarsen ratio", "0.0", 'd');
  d(C, "Number of initial global refinments", "4", 'i');
  d(C, "Number of n-cycle in AdditionalData", "4", 'i');
  d(C, "maximum linear iteration number", "10000", 'i');
  d(C, "Using dampped Newton iteration", "true", 'b');
  d(C, "primary step length of dampped newton iteration", "0.83", 'd');
  // the build's own keys (compile-time choices in the reference)
  d(X, "geometry", "retangle", 's');                 // retangle | cube (makegrid_* variants)
  d(X, "boundary ids", "1 1 1 1 4 4", 's');          // faces -x +x -y +y -z +z (5 5: periodic axis)
  d(X, "finite element degree", "1", 'i');
  d(X, "initial guess", "B", 's');                   // B | A | BnA | BinA | BnB | PDW
  d(X, "random seed", "0", 'i');
  d(X, "noise amplitude", "0.05", 'd');
  d(X, "mixed precision", "false", 'b');
  d(X, "matrix free", "true", 'b');                  // matrix-free Jacobian where it applies (glh_set_matrix_free)
  d(X, "GMRES restart", "100", 'i');
  d(X, "output directory", "", 's');                 // output_results per iteration (empty: none)
  d(X, "verbose", "false", 'b');                     // progress lines of run() on stdout
}

// mkdir that accepts an existing directory
bool make_dir(const std::string &path)
{
  if (::mkdir(path.c_str(), 0755) == 0) return true;
  struct stat sb;
  return errno == EEXIST && ::stat(path.c_str(), &sb) == 0 && S_ISDIR(sb.st_mode);
}

bool check_value(const Entry &e, const std::string &v, std::string &why)
{
  if (e.type == 's') return true;
  if (e.type == 'b') {
    if (v == "true" || v == "false") return true;
    why = "expected true or false";
    return false;
  }
  char *end = nullptr;
  if (e.type == 'i') {
    const long x = std::strtol(v.c_str(), &end, 10);
    if (!end || *end || v.empty()) { why = "expected an integer"; return false; }
    if (e.nonneg && x < 0) { why = "expected a non-negative integer"; return false; }
    return true;
  }
  const double x = std::strtod(v.c_str(), &end);
  if (!end || *end || v.empty() || !std::isfinite(x)) { why = "expected a number"; return false; }
  if (e.nonneg && x < 0) { why = "expected a non-negative number"; return false; }
  return true;
}

}  // namespace

struct glh_prm {
  std::map<std::string, Entry> e;
  const Entry *find(const char *sub, const char *key) const
  {
    auto it = e.find(std::string(sub) + "/" + key);
    return it == e.end() ? nullptr : &it->second;
  }
  double d(const char *sub, const char *key) const { return std::strtod(find(sub, key)->value.c_str(), nullptr); }
  long i(const char *sub, const char *key) const { return std::strtol(find(sub, key)->value.c_str(), nullptr, 10); }
  bool b(const char *sub, const char *key) const { return find(sub, key)->value == "true"; }
  std::string s(const char *sub, const char *key) const { return find(sub, key)->value; }
};

extern "C" {

int glh_prm_parse(const char *text, glh_prm **out, char *err, int err_len)
{
  auto report = [&](const std::string &msg) {
    if (err && err_len > 0) {
      std::strncpy(err, msg.c_str(), (size_t)err_len - 1);
      err[err_len - 1] = 0;
    }
    return GLH_ERR_ARG;
  };
  if (!text || !out) return report("null argument");
  *out = nullptr;
  auto *p = new glh_prm();
  declare(p->e);
  std::istringstream in(text);
  std::string line;
  std::vector<std::string> path;
  int ln = 0;
  while (std::getline(in, line)) {
    ++ln;
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line = line.substr(0, hash);
    line = trim(line);
    if (line.empty()) continue;
    auto fail = [&](const std::string &msg) {
      delete p;
      return report("line " + std::to_string(ln) + ": " + msg);
    };
    if (line.compare(0, 11, "subsection ") == 0) {
      path.push_back(trim(line.substr(11)));
    } else if (line == "end") {
      if (path.empty()) return fail("'end' without a subsection");
      path.pop_back();
    } else if (line.compare(0, 4, "set ") == 0) {
      const size_t eq = line.find('=');
      if (eq == std::string::npos) return fail("'set' without '='");
      const std::string key = trim(line.substr(4, eq - 4)), val = trim(line.substr(eq + 1));
      std::string sub;
      for (size_t k = 0; k < path.size(); ++k) sub += (k ? "/" : "") + path[k];
      auto it = p->e.find(sub + "/" + key);
      if (it == p->e.end()) return fail("undeclared entry '" + key + "' in subsection '" + sub + "'");
      std::string why;
      if (!check_value(it->second, val, why)) return fail("entry '" + key + "': " + why + " (got '" + val + "')");
      it->second.value = val;
    } else {
      return fail("expected 'subsection', 'set' or 'end'");
    }
  }
  if (!path.empty()) {
    delete p;
    return report("subsection '" + path.back() + "' is not closed");
  }
  *out = p;
  return GLH_OK;
}

int glh_prm_destroy(glh_prm *p)
{
  delete p;
  return GLH_OK;
}

int glh_prm_get(const glh_prm *p, const char *subsection, const char *key, char *value, int len)
{
  if (!p || !subsection || !key || !value || len < 1) return GLH_ERR_ARG;
  const Entry *e = p->find(subsection, key);
  if (!e) return GLH_ERR_ARG;
  std::strncpy(value, e->value.c_str(), (size_t)len - 1);
  value[len - 1] = 0;
  return GLH_OK;
}

int glh_prm_get_double(const glh_prm *p, const char *subsection, const char *key, double *v)
{
  if (!p || !subsection || !key || !v) return GLH_ERR_ARG;
  const Entry *e = p->find(subsection, key);
  if (!e || e->type == 'b' || e->type == 's') return GLH_ERR_ARG;
  *v = std::strtod(e->value.c_str(), nullptr);
  return GLH_OK;
}

int glh_run_prm(glh_ctx *ctx, const glh_prm *p, int max_rows, double *rows, int *n_rows)
{
  if (!ctx || !p || !rows || !n_rows || max_rows < 1) return GLH_ERR_ARG;
  *n_rows = 0;
  const char *P = "physical parameters", *C = "control parameters", *X = "glhip";
  const double pres = p->d(P, "pressure in bar"), t = p->d(P, "t_reduced"), bt = p->d(P, "AdGR diffuse length");
  const int scc = p->b(P, "trun on Strong Coupling Correction");
  const int n_cycles = (int)p->i(C, "Number of refinements"), n_iter = (int)p->i(C, "Number of interations");
  const double conv = p->d(C, "converge accuracy"), frac = p->d(C, "adaptive refinment ratio");
  const double cfrac = p->d(C, "adaptive coarsen ratio");
  if (frac + cfrac > 1.0) return GLH_ERR_ARG;
  const int maxit = (int)std::max(1L, p->i(C, "maximum linear iteration number"));
  const int damped = p->b(C, "Using dampped Newton iteration");
  const double step = p->d(C, "primary step length of dampped newton iteration");
  const int n_glob = (int)p->i(C, "Number of initial global refinments");
  const int deg = (int)p->i(X, "finite element degree");
  const int fp32 = p->b(X, "mixed precision");
  if (const int rmf = glh_set_matrix_free(ctx, p->b(X, "matrix free"))) return rmf;
  const int restart = (int)std::max(1L, std::min(126L, p->i(X, "GMRES restart")));
  if (deg != 1 && deg != 2) return GLH_ERR_ARG;
  const std::string geo = p->s(X, "geometry"), guess = p->s(X, "initial guess");
  if ((geo != "retangle" && geo != "cube") || (guess != "B" && guess != "A" && guess != "BnA" && guess != "BinA" && guess != "BnB" && guess != "PDW"))
    return GLH_ERR_ARG;
  if (n_glob < 0 || n_glob > 6 || n_cycles < 0 || n_cycles > 8) return GLH_ERR_ARG;
  double thr[4], tol[5];
  bool glob[5] = {false, false, false, false, false};
  for (int k = 0; k < 4; ++k) thr[k] = p->d(C, ("Cycle " + std::to_string(k) + " refinement threshold").c_str());
  for (int k = 0; k < 5; ++k) tol[k] = p->d(C, ("Cycle " + std::to_string(k) + " linear solver tol").c_str());
  for (int k = 1; k < 5; ++k) glob[k] = p->b(C, ("Cycle " + std::to_string(k) + " do global refinement").c_str());
  // make_grid: hyper_rectangle / hyper_cube refined n_glob times
  double lo[3], hi[3];
  if (p->s(X, "geometry") == "cube") {
    const double L = p->d(C, "cube half side length");
    for (int k = 0; k < 3; ++k) { lo[k] = -L; hi[k] = L; }
  } else {
    const double L[3] = {p->d(C, "half x length of retangle"), p->d(C, "half y length of retangle"),
                         p->d(C, "half z length of retangle")};
    for (int k = 0; k < 3; ++k) { lo[k] = -L[k]; hi[k] = L[k]; }
  }
  int sid[6];
  {
    std::istringstream ss(p->s(X, "boundary ids"));
    for (int k = 0; k < 6; ++k)
      if (!(ss >> sid[k]) || sid[k] < 0 || sid[k] > 5) return GLH_ERR_ARG;  // 5: periodic pair
  }
  double mp[12];
  if (glh_matep(pres, t, scc, mp) != GLH_OK) return GLH_ERR_RANGE;
  const double K[3] = {0.42072, 0.42072, 0.42072};  // femgl.h:320-322
  int rc = glh_set_material(ctx, mp[0], mp + 1, K, bt);
  if (rc) return rc;
  // The reference refines one coarse cell n_glob times, so every initial cell
  // has a parent that the coarsen ratio can restore. Here the forest starts
  // from 2 roots per axis (a periodic axis needs two) refined n_glob - 1
  // times: the same cells, coarsenable down to the roots.
  const int nroot = n_glob > 0 ? 2 : 1;
  glh_forest *f = nullptr;
  if ((rc = glh_forest_create(nroot, nroot, nroot, lo, hi, sid, &f)) != GLH_OK) return rc;
  if ((rc = glh_forest_set_degree(f, deg)) != GLH_OK) {
    glh_forest_destroy(f);
    return rc;
  }
  for (int g = 1; g < n_glob && rc == GLH_OK; ++g) {
    std::vector<uint8_t> all(glh_forest_n_cells(f), 1);
    glh_forest *fn = nullptr;
    if ((rc = glh_forest_refine(f, all.data(), &fn)) == GLH_OK) {
      glh_forest_destroy(f);
      f = fn;
    }
  }
  if (rc) {
    glh_forest_destroy(f);
    return rc;
  }
  glh_mesh *m = nullptr;
  std::vector<double> U;
  double res = 0.0;
  for (int cycle = 0; cycle <= n_cycles && rc == GLH_OK; ++cycle) {
    if (cycle > 0) {  // refine_grid (refine.cc:111-175)
      const int nc = (int)glh_forest_n_cells(f);
      std::vector<uint8_t> fl(nc, 1);
      if (!glob[std::min(cycle, 4)]) {
        std::vector<float> eta(nc);
        if ((rc = glh_kelly(f, m, U.data(), eta.data())) ||
            (rc = glh_flag_refine_coarsen(nc, eta.data(), frac, cfrac, fl.data())))
          break;
      }
      glh_forest *fn = nullptr;
      glh_mesh *mn = nullptr;
      if ((rc = glh_forest_refine(f, fl.data(), &fn)) || (rc = glh_forest_mesh(fn, &mn))) break;
      std::vector<double> Un((size_t)mn->m.n_nodes * 18);
      rc = glh_forest_transfer(f, m, U.data(), fn, mn, Un.data());
      glh_mesh_destroy(m);
      glh_forest_destroy(f);
      f = fn;
      m = mn;
      U.swap(Un);
      if (rc) break;
    } else {  // make_grid + setup_system with the configured initial guess
      if ((rc = glh_forest_mesh(f, &m))) break;
      const int nn = m->m.n_nodes;
      U.assign((size_t)nn * 18, 0.0);
      const std::string g = p->s(X, "initial guess");
      const double amp = p->d(X, "noise amplitude");
      const uint64_t seed = (uint64_t)p->i(X, "random seed");
      if (g == "BnA")
        rc = glh_initial_guess_bna(nn, m->m.node_xyz.data(), p->d(C, "A-phase block range ratio") * hi[2], mp[8],
                                   mp[9], amp, seed, U.data());
      else if (g == "BinA") {  // setup_uniform_BinA-bubble-like-configuration.cc:214-235
        const double par[3] = {p->d(C, "B-phase ball radius ratio") * p->d(C, "cube half side length"), mp[8], mp[9]};
        rc = glh_initial_guess_field(GLH_GUESS_BINA, nn, m->m.node_xyz.data(), par, amp, seed, U.data());
      } else if (g == "BnB") {  // setup_uniform_BnB-polkka-dot-like-configuration.cc:289-313
        const double par[1] = {p->d(C, "B-phase inner plate radius ratio") * p->d(C, "half x length of retangle")};
        rc = glh_initial_guess_field(GLH_GUESS_BNB, nn, m->m.node_xyz.data(), par, amp, seed, U.data());
      } else if (g == "PDW") {  // setup_weak-coupling-PDW-configuration.cc:285-315
        const double par[3] = {p->d(P, "amplitude u parameter"), p->d(C, "half z length of retangle"), mp[9]};
        rc = glh_initial_guess_field(GLH_GUESS_PDW, nn, m->m.node_xyz.data(), par, amp, seed, U.data());
      } else
        rc = glh_initial_guess_host(nn, g == "A" ? mp[8] : mp[9], g == "A" ? 1 : 0, amp, seed, U.data());
      if (rc) break;
    }
    if ((rc = glh_upload_mesh(ctx, m)) || (rc = glh_set_solution(ctx, U.data()))) break;
    // the iteration loop of run.cc:202-249
    double res_last = 0.0;
    int it = 0, lin_total = 0;
    const std::string outdir = p->s(X, "output directory");
    const bool verbose = p->b(X, "verbose");
    std::string cycdir;
    if (!outdir.empty()) {  // "./refine-cycle_<cycle>/" of run.cc:222
      cycdir = outdir + "/refine-cycle_" + std::to_string(cycle);
      if (!make_dir(outdir) || !make_dir(cycdir)) { rc = GLH_ERR_ARG; break; }
    }
    for (it = 0; it <= n_iter; ++it) {
      double st[2] = {0, 0}, out[4];
      if (verbose) std::printf("Refinement cycle : %d, iteration_loop: %d\n", cycle, it);
      if ((rc = glh_assemble(ctx, GLH_ASM_J | GLH_ASM_R | (fp32 ? GLH_ASM_FP32 : 0)))) break;
      rc = glh_solve(ctx, tol[std::min(cycle, 4)], maxit, restart, st);
      if (rc) break;  // SolverControl::NoConvergence ends the run, as in the reference
      lin_total += (int)st[0];
      if ((rc = glh_newton_step(ctx, step, damped, out))) break;
      res = out[1];
      if (verbose) {
        std::printf(" linear solve: %d iterations (tol %g), Newton step: ||R|| %.6e\n", (int)st[0],
                    tol[std::min(cycle, 4)], res);
        std::fflush(stdout);
      }
      if (!cycdir.empty()) {  // output_results (io.cc:105-168), counter = iteration_loop
        std::vector<double> dU(U.size());
        if ((rc = glh_get_solution(ctx, U.data())) || (rc = glh_get_update(ctx, dU.data())) ||
            (rc = glh_write_vtu(m, U.data(), dU.data(), cycdir.c_str(), it)))
          break;
      }
      // stuck: refine (the cycle test first, so no threshold beyond the four is read)
      if (cycle < n_cycles && res > conv && std::fabs(res - res_last) < thr[std::min(cycle, 3)]) break;
      if (res <= conv) break;
      res_last = res;
    }
    if (rc) break;
    if ((rc = glh_get_solution(ctx, U.data()))) break;
    if (*n_rows < max_rows) {
      double F = 0;
      glh_free_energy(ctx, &F);
      double *r = rows + (size_t)(*n_rows) * 6;
      r[0] = cycle; r[1] = (double)m->m.n_cells; r[2] = (double)m->m.n_nodes * 18; r[3] = std::min(it, n_iter) + 1;
      r[4] = res; r[5] = F;
      ++*n_rows;
    }
    if (res <= conv) break;  // run.cc:252-253
  }
  if (m) glh_mesh_destroy(m);
  glh_forest_destroy(f);
  return rc;
}

}  // extern "C"
