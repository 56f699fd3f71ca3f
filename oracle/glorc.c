/*
 * glorc.c -- FP64 CPU oracle for the femgl Ginzburg-Landau Newton step.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links or calls
 * this file; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load it, and only as the checker or the CPU baseline.
 *
 * It restates, in plain C, the arithmetic of the reference femgl solver
 * (VerHem/verkko-Hem-repo).  Every function cites the reference file:line it
 * follows.  Two forms of the cell assembly exist:
 *   - "faithful": the literal per-(i, j, q) loop of 3x3 matrix products of
 *     femgl/src/assemble.cc:188-276 with the term functions of
 *     femgl/src/cell_mat_vec/ files, used on meshes of a few cells;
 *   - "factorized": the same bilinear forms evaluated once per quadrature
 *     point as an 18x18 Hessian / 18-vector and contracted with the scalar
 *     shape functions (exact for FESystem(FE_Q(k),18), whose shape functions
 *     are one-hot in the component).
 * The faithful form pins the factorized one; finite differences of the free
 * energy pin both (tests/test_oracle_*.py).
 *
 * Local ordering inside a cell: dof i = a*18 + c, a = ix + n1*(iy + n1*iz)
 * lexicographic over the (k+1)^3 Lagrange support points, c = component
 * (c < 9: u_{c/3, c%3}; c >= 9: v_{(c-9)/3, (c-9)%3}),
 * femgl/src/s_vector2matrix.cc:159-160.
 *
 * Parity status: every function here is pinned by a test (closed forms,
 * finite differences, brute force on tiny inputs), except the deal.II
 * global DoF enumeration and the constrained-row diagonal value, which
 * only a deal.II build could fix ("parity unpinned" for those two items;
 * the oracle uses deal.II's documented |local diagonal| rule).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NC 18

/* ------------------------------------------------------------------ */
/* matep: matep/src/matep.cc, matep/inc/matep.h                         */
/* ------------------------------------------------------------------ */

/* float literals stored in doubles (matep.h:102-103, matep.cc:46-52) */
static double mp_pi(void) { return (double)3.14159265358979323846264338328f; }
static double mp_zeta3(void) { return (double)1.2020569031595942f; }
/* c_betai = (7.0f*zeta3)/(80.0f*pi*pi), matep.cc:52 (double arithmetic) */
static double mp_cbetai(void)
{
  double pi = mp_pi();
  return (7.0 * mp_zeta3()) / ((80.0 * pi) * pi);
}

/* strong-coupling tables, matep.cc:59-63 (0..34 bar, 2-bar steps) */
static const double MP_C[5][18] = {
  {-0.0098, -0.0127, -0.0155, -0.0181, -0.0207, -0.0231, -0.0254, -0.0275, -0.0295,
   -0.0314, -0.0330, -0.0345, -0.0358, -0.0370, -0.0381, -0.0391, -0.0402, -0.0413},
  {-0.0419, -0.0490, -0.0562, -0.0636, -0.0711, -0.0786, -0.0861, -0.0936, -0.1011,
   -0.1086, -0.1160, -0.1233, -0.1306, -0.1378, -0.1448, -0.1517, -0.1583, -0.1645},
  {-0.0132, -0.0161, -0.0184, -0.0202, -0.0216, -0.0226, -0.0233, -0.0239, -0.0243,
   -0.0247, -0.0249, -0.0252, -0.0255, -0.0258, -0.0262, -0.0265, -0.0267, -0.0268},
  {-0.0047, -0.0276, -0.0514, -0.0760, -0.1010, -0.1260, -0.1508, -0.1751, -0.1985,
   -0.2208, -0.2419, -0.2614, -0.2795, -0.2961, -0.3114, -0.3255, -0.3388, -0.3518},
  {-0.0899, -0.1277, -0.1602, -0.1880, -0.2119, -0.2324, -0.2503, -0.2660, -0.2801,
   -0.2930, -0.3051, -0.3167, -0.3280, -0.3392, -0.3502, -0.3611, -0.3717, -0.3815}};

/* lininterp, matep.cc:365-406: left-closed 2-bar bins, result rounded
 * through a float.  p outside [0,34) is undefined behaviour in the
 * reference; the oracle returns NaN there. */
static double mp_lininterp(const double *c, double p)
{
  if (!(p >= 0.0 && p < 34.0)) return NAN;
  int k = (int)floor(p / 2.0);
  if (k > 16) k = 16;
  /* make the bin test identical to the reference's chained comparisons */
  while (k > 0 && p < 2.0 * k) --k;
  while (k < 16 && p >= 2.0 * (k + 1)) ++k;
  float pk = (float)(2 * k);
  float fp = (float)(((c[k + 1] - c[k]) / 2.0) * (p - (double)pk) + c[k]);
  return (double)fp;
}

static const double MP_B[5] = {-1.0, 2.0, 2.0, 2.0, -2.0}; /* matep.cc:149-201 */

/* out: [0]=alpha, [1..5]=beta1..5, [6]=beta_A, [7]=beta_B, [8]=gap_A,
 *      [9]=gap_B, [10]=f_A, [11]=f_B.  Returns 0, or -1 if p is outside
 *      the table (reference UB). */
int orc_matep(double p, double t, int scc, double *out)
{
  double cb = mp_cbetai();
  double b[5];
  out[0] = 1.0 * (t - 1.0);                                /* matep.cc:141 */
  for (int n = 0; n < 5; ++n) {
    if (scc) b[n] = cb * (MP_B[n] + t * mp_lininterp(MP_C[n], p));
    else     b[n] = cb * MP_B[n];
    out[1 + n] = b[n];
  }
  double bA = (b[1] + b[3]) + b[4];                         /* matep.cc:213-216 */
  double bB = (b[0] + b[1]) + (double)(1.f / 3.f) * ((b[2] + b[3]) + b[4]); /* :218-221 */
  out[6] = bA;
  out[7] = bB;
  double a = out[0];
  out[8] = (t <= 1.0) ? sqrt(-a / (2.0 * bA)) : 0.0;        /* :224-234 */
  out[9] = (t <= 1.0) ? sqrt(-a / (2.0 * bB)) : 0.0;        /* :237-248 */
  out[10] = (t <= 1.0) ? (-0.25 * pow(a, 2.0)) / bA : 0.0;  /* :310-319 */
  out[11] = (t <= 1.0) ? (-0.25 * pow(a, 2.0)) / bB : 0.0;  /* :322-333 */
  return (scc && isnan(b[0])) ? -1 : 0;
}

/* ------------------------------------------------------------------ */
/* 3x3 helpers mirroring deal.II FullMatrix mTmult / Tmmult / mmult    */
/* ------------------------------------------------------------------ */
static void mT(const double *A, const double *B, double *C) /* C = A B^T */
{
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += A[3 * i + k] * B[3 * j + k];
      C[3 * i + j] = s;
    }
}
static void Tm(const double *A, const double *B, double *C) /* C = A^T B */
{
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += A[3 * k + i] * B[3 * k + j];
      C[3 * i + j] = s;
    }
}
static void mm(const double *A, const double *B, double *C) /* C = A B */
{
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += A[3 * i + k] * B[3 * k + j];
      C[3 * i + j] = s;
    }
}
static double tr(const double *A) { return A[0] + A[4] + A[8]; }
static double trmm(const double *A, const double *B) /* Tr(A B) */
{
  double C[9];
  mm(A, B, C);
  return tr(C);
}
static double trmT(const double *A, const double *B) /* Tr(A B^T) = A:B */
{
  double s = 0;
  for (int i = 0; i < 9; ++i) s += A[i] * B[i];
  return s;
}

/* ------------------------------------------------------------------ */
/* Literal term functions, femgl/src/cell_mat_vec/ (one file per term)   */
/* u, v: the iterate at q; pi/qi: phi_u/phi_v of dof i; pj/qj of dof j */
/* ------------------------------------------------------------------ */

/* cell_mat_lhs_alpha.cc:109-143 */
static double lit_mat_alpha(const double *pi, const double *pj, const double *qi, const double *qj)
{
  return trmT(pi, pj) + trmT(qi, qj);
}

/* cell_mat_lhs_K1.cc:109-141; g*[k] are 3x3 gradient matrices */
static double lit_mat_K1(const double *gpi, const double *gqi, const double *gpj, const double *gqj)
{
  double s = 0;
  for (int k = 0; k < 3; ++k) s += trmT(gpi + 9 * k, gpj + 9 * k) + trmT(gqi + 9 * k, gqj + 9 * k);
  return s;
}

/* orbital divergence d(mu) = sum_z grad[z](mu, z): cell_mat_lhs_K2K3.cc:150-178 */
static void divo(const double *g, double *d)
{
  for (int m = 0; m < 3; ++m) {
    d[m] = 0;
    for (int z = 0; z < 3; ++z) d[m] += g[9 * z + 3 * m + z];
  }
}

/* cell_mat_lhs_K2K3.cc:109-184 */
static double lit_mat_K2K3(const double *gpi, const double *gqi, const double *gpj, const double *gqj)
{
  double a[3], b[3], c[3], d[3];
  divo(gpi, a); divo(gpj, b); divo(gqi, c); divo(gqj, d);
  return (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]) + (c[0] * d[0] + c[1] * d[1] + c[2] * d[2]);
}

/* cell_mat_lhs_beta1.cc:109-198 */
static double lit_mat_b1(const double *u, const double *v, const double *pi, const double *pj,
                         const double *qi, const double *qj)
{
  double uu = trmT(u, u), vv = trmT(v, v), uv = trmT(u, v);
  double pp = trmT(pi, pj), qq = trmT(qi, qj), pq = trmT(pi, qj), qp = trmT(qi, pj);
  double u_pi = trmT(u, pi), u_pj = trmT(u, pj), v_qi = trmT(v, qi), v_qj = trmT(v, qj);
  double u_qi = trmT(u, qi), u_qj = trmT(u, qj), v_pi = trmT(v, pi), v_pj = trmT(v, pj);
  return uu * (pp - qq) + 2.0 * ((v_pi + u_qi) * v_pj) + vv * (qq - pp) +
         2.0 * ((v_qi - u_pi) * v_qj) + 2.0 * (uv * (qp + pq)) +
         2.0 * ((v_pi * u_qj) - (v_qi * u_pj)) + 2.0 * ((u_pi * u_pj) + (u_qi * u_qj));
}

/* cell_mat_lhs_beta2.cc:109-168 */
static double lit_mat_b2(const double *u, const double *v, const double *pi, const double *pj,
                         const double *qi, const double *qj)
{
  return (trmT(u, u) + trmT(v, v)) * (trmT(pi, pj) + trmT(qi, qj)) +
         2.0 * ((trmT(u, pi) + trmT(v, qi)) * (trmT(u, pj) + trmT(v, qj)));
}

/* cell_mat_lhs_beta3.cc:109-335 (trace polynomials I..IV, :262-334) */
static double lit_mat_b3(const double *u, const double *v, const double *pi, const double *pj,
                         const double *qi, const double *qj)
{
  double pu_pu[9], pv_pv[9], pu_pvt[9], pv_put[9], put_pu[9], pvt_pv[9], put_pv[9], pvt_pu[9];
  double u0u0t[9], u0v0t[9], v0u0t[9], v0tv0[9], v0tu0[9];
  double u_pit[9], u_pjt[9], v_qit[9], v_qjt[9], u_qit[9], u_qjt[9], v_pit[9], v_pjt[9];
  double vt_pi[9], vt_qi[9], pj_ut[9], qj_ut[9], pjt_v[9], qjt_v[9];
  double X[9];
  mT(pi, pj, pu_pu); mT(qi, qj, pv_pv); mT(pi, qj, pu_pvt); mT(qi, pj, pv_put);
  Tm(pi, pj, put_pu); Tm(qi, qj, pvt_pv); Tm(pi, qj, put_pv); Tm(qi, pj, pvt_pu);
  mT(u, u, u0u0t); mT(u, v, u0v0t); mT(v, u, v0u0t); Tm(v, v, v0tv0); Tm(v, u, v0tu0);
  mT(u, pi, u_pit); mT(u, pj, u_pjt); mT(v, qi, v_qit); mT(v, qj, v_qjt);
  mT(u, qi, u_qit); mT(u, qj, u_qjt); mT(v, pi, v_pit); mT(v, pj, v_pjt);
  Tm(v, pi, vt_pi); Tm(v, qi, vt_qi);
  mT(pj, u, pj_ut); mT(qj, u, qj_ut);
  Tm(pj, v, pjt_v); Tm(qj, v, qjt_v);
  (void)X;
  double I = trmm(u_pit, u_pjt) + trmm(v_pit, v_pjt) + trmm(v_pit, u_qjt) - trmm(u_pit, v_qjt);
  double II = -trmm(v_qit, u_pjt) + trmm(u_qit, v_pjt) + trmm(u_qit, u_qjt) + trmm(v_qit, v_qjt);
  double III = trmm(u_pit, pj_ut) + trmm(u0v0t, pu_pvt) + trmm(u0v0t, pv_put) - trmm(v_qit, pj_ut) +
               trmm(u0u0t, pu_pu) - trmm(u0u0t, pv_pv) + trmm(v_pit, qj_ut) + trmm(u_qit, qj_ut);
  double IV = trmm(v0tv0, put_pu) + trmm(v0u0t, pu_pvt) + trmm(v0u0t, pv_put) + trmm(v0tu0, pvt_pu) -
              trmm(vt_pi, pjt_v) + trmm(vt_qi, qjt_v) - trmm(v0tu0, put_pv) + trmm(v0tv0, pvt_pv);
  return I + II + III + IV;
}

/* cell_mat_lhs_beta4.cc:109-292 */
static double lit_mat_b4(const double *u, const double *v, const double *pi, const double *pj,
                         const double *qi, const double *qj)
{
  double t1[9], t2[9];
  double ms1[9], ms2[9], ms3[9], ms4[9], ms5[9], ms6[9];
  double u0u0t[9], u0v0t[9], v0v0t[9], v0u0t[9];
  double u_pjt[9], u_qjt[9], v_pjt[9], v_qjt[9];
  mT(u, u, u0u0t); mT(u, v, u0v0t); mT(v, v, v0v0t); mT(v, u, v0u0t);
  mT(u, pj, u_pjt); mT(u, qj, u_qjt); mT(v, pj, v_pjt); mT(v, qj, v_qjt);
  mT(pi, u, t1); mT(qi, v, t2);
  for (int k = 0; k < 9; ++k) ms1[k] = t1[k] + t2[k];           /* phi_ui u^T + phi_vi v^T */
  mT(pj, u, t1); mT(qj, v, t2);
  for (int k = 0; k < 9; ++k) ms2[k] = t1[k] + t2[k];
  mT(pi, v, t1); mT(qi, u, t2);
  for (int k = 0; k < 9; ++k) ms3[k] = t1[k] - t2[k];           /* phi_ui v^T - phi_vi u^T */
  mT(qj, u, t1); mT(pj, v, t2);
  for (int k = 0; k < 9; ++k) ms4[k] = t1[k] - t2[k];           /* phi_vj u^T - phi_uj v^T */
  mT(qi, pj, t1); mT(pi, qj, t2);
  for (int k = 0; k < 9; ++k) ms5[k] = t1[k] - t2[k];
  mT(pi, pj, t1); mT(qi, qj, t2);
  for (int k = 0; k < 9; ++k) ms6[k] = t1[k] + t2[k];
  double I = trmm(ms1, ms2) + trmm(ms3, ms4);
  double II = trmm(ms1, u_pjt) + trmm(ms5, u0v0t) + trmm(ms6, u0u0t) - trmm(ms3, u_qjt);
  double III = trmm(ms3, v_pjt) - trmm(ms5, v0u0t) + trmm(ms6, v0v0t) + trmm(ms1, v_qjt);
  return I + II + III;
}

/* cell_mat_lhs_beta5.cc:109-342 */
static double lit_mat_b5(const double *u, const double *v, const double *pi, const double *pj,
                         const double *qi, const double *qj)
{
  double t1[9], t2[9], t3[9], t4[9];
  double ms1[9], ms2[9], ms3[9], ms4[9], ms5[9], ms6[9], ms7[9], ms8[9], ms9[9];
  double u0v0t[9], u0tv0[9], u_pit[9], v_pit[9], ut_qi[9], v_qit[9], u_qit[9];
  double a[9], b[9];
  mT(u, u, t1); mT(v, v, t2);
  for (int k = 0; k < 9; ++k) ms1[k] = t1[k] + t2[k];
  mT(pi, pj, t1); mT(qi, qj, t2);
  for (int k = 0; k < 9; ++k) ms2[k] = t1[k] + t2[k];
  mT(u, pj, t1); mT(pj, u, t2); mT(v, qj, t3); mT(qj, v, t4);
  for (int k = 0; k < 9; ++k) ms3[k] = t1[k] + t2[k] + t3[k] + t4[k];
  mT(v, pj, t1); mT(pj, v, t2); mT(u, qj, t3);
  for (int k = 0; k < 9; ++k) ms4[k] = t1[k] - t2[k] - t3[k];
  mT(pi, qj, t1); mT(qi, pj, t2);
  for (int k = 0; k < 9; ++k) ms5[k] = t1[k] - t2[k];
  Tm(pj, v, t1); Tm(u, qj, t2);
  for (int k = 0; k < 9; ++k) ms6[k] = t1[k] + t2[k];
  mT(u, pj, t1); mT(qj, v, t2);
  for (int k = 0; k < 9; ++k) ms7[k] = t1[k] + t2[k];
  mT(v, pj, t1); mT(qj, u, t2); mT(pj, v, t3);
  for (int k = 0; k < 9; ++k) ms8[k] = t1[k] + t2[k] - t3[k];
  Tm(qi, pj, t1); Tm(pi, qj, t2);
  for (int k = 0; k < 9; ++k) ms9[k] = t1[k] + t2[k];
  mT(u, v, u0v0t); Tm(u, v, u0tv0);
  mT(u, pi, u_pit); mT(v, pi, v_pit); Tm(u, qi, ut_qi); mT(v, qi, v_qit); mT(u, qi, u_qit);
  double I = trmm(ms1, ms2) + trmm(u_pit, ms3) + trmm(v_pit, ms4) + trmm(u0v0t, ms5) +
             trmm(ut_qi, ms6) + trmm(v_qit, ms7) + trmm(u0tv0, ms9);
  Tm(v, qi, a); Tm(v, qj, b);
  I += trmm(a, b);                                   /* mm1 = (v^T phi_vi)(v^T phi_vj) */
  double II = trmm(u_qit, ms8);
  Tm(u, pi, a); Tm(qj, v, b);
  double mm2 = trmm(a, b);                           /* (u^T phi_ui)(phi_vj^T v) */
  return I - II - mm2;
}

/* cell_vec_rhs_alpha.cc:109-139 */
static double lit_vec_alpha(const double *pi, const double *qi, const double *u, const double *v)
{
  return trmT(u, pi) + trmT(v, qi);
}

/* cell_vec_rhs_K1.cc:109-139 */
static double lit_vec_K1(const double *gu, const double *gv, const double *gpi, const double *gqi)
{
  double s = 0;
  for (int k = 0; k < 3; ++k) s += trmT(gu + 9 * k, gpi + 9 * k) + trmT(gv + 9 * k, gqi + 9 * k);
  return s;
}

/* cell_vec_rhs_K2K3.cc:109-170 */
static double lit_vec_K2K3(const double *gpi, const double *gqi, const double *gu, const double *gv)
{
  double a[3], b[3], c[3], d[3];
  divo(gu, a); divo(gpi, b); divo(gv, c); divo(gqi, d);
  return (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]) + (c[0] * d[0] + c[1] * d[1] + c[2] * d[2]);
}

/* cell_vec_rhs_beta1.cc:109-165 */
static double lit_vec_b1(const double *u, const double *v, const double *pi, const double *qi)
{
  return (trmT(u, u) - trmT(v, v)) * (trmT(u, pi) - trmT(v, qi)) +
         2.0 * trmT(u, v) * (trmT(u, qi) + trmT(v, pi));
}

/* cell_vec_rhs_beta2.cc:109-146 */
static double lit_vec_b2(const double *u, const double *v, const double *pi, const double *qi)
{
  return (trmT(u, u) + trmT(v, v)) * (trmT(u, pi) + trmT(v, qi));
}

/* cell_vec_rhs_beta3.cc:109-204 */
static double lit_vec_b3(const double *u, const double *v, const double *pi, const double *qi)
{
  double utu[9], uvt[9], utv[9], vtv[9], ut_pi[9], ut_qi[9], vt_pi[9], vt_qi[9], v_pit[9];
  double pit_v[9], qit_u[9];
  mT(u, v, uvt); Tm(u, u, utu); Tm(u, v, utv); Tm(v, v, vtv);
  Tm(u, pi, ut_pi); Tm(u, qi, ut_qi); Tm(v, pi, vt_pi); Tm(v, qi, vt_qi);
  mT(v, pi, v_pit); Tm(pi, v, pit_v); Tm(qi, u, qit_u);
  return trmm(utu, ut_pi) + trmm(uvt, v_pit) + trmm(utv, pit_v) - trmm(utv, vt_pi) -
         trmm(utv, qit_u) + trmm(utv, ut_qi) + trmm(utu, vt_qi) + trmm(vtv, vt_qi);
}

/* cell_vec_rhs_beta4.cc:109-210 */
static double lit_vec_b4(const double *u, const double *v, const double *pi, const double *qi)
{
  double uut[9], uvt[9], utu[9], utv[9], vtv[9], t1[9], t2[9];
  double ms1[9], ms2[9], ms3[9], pi_vt[9], qi_vt[9];
  mT(u, u, uut); mT(u, v, uvt); Tm(u, u, utu); Tm(u, v, utv); Tm(v, v, vtv);
  for (int k = 0; k < 9; ++k) ms1[k] = utu[k] + vtv[k];
  Tm(u, pi, t1); Tm(v, qi, t2);
  for (int k = 0; k < 9; ++k) ms2[k] = t1[k] + t2[k];
  Tm(v, pi, t1); Tm(u, qi, t2);
  for (int k = 0; k < 9; ++k) ms3[k] = t1[k] - t2[k];
  mT(pi, v, pi_vt); mT(qi, v, qi_vt);
  return trmm(ms1, ms2) + trmm(utv, ms3) + trmm(uut, qi_vt) - trmm(uvt, pi_vt);
}

/* cell_vec_rhs_beta5.cc:109-216 */
static double lit_vec_b5(const double *u, const double *v, const double *pi, const double *qi)
{
  double uvt[9], utu[9], utv[9], vtv[9], t1[9], t2[9], t3[9], t4[9], ms1[9], ms2[9];
  double v_pit[9], vt_qi[9];
  mT(u, v, uvt); Tm(u, u, utu); Tm(u, v, utv); Tm(v, v, vtv);
  Tm(u, pi, t1); Tm(v, qi, t2);
  for (int k = 0; k < 9; ++k) ms1[k] = t1[k] - t2[k];
  Tm(pi, v, t1); Tm(v, pi, t2); Tm(qi, u, t3); Tm(u, qi, t4);
  for (int k = 0; k < 9; ++k) ms2[k] = t1[k] + t2[k] + t3[k] + t4[k];
  mT(v, pi, v_pit); Tm(v, qi, vt_qi);
  return trmm(utu, ms1) + trmm(utv, ms2) - trmm(uvt, v_pit) + trmm(vtv, vt_qi);
}

/* coefficient vector used throughout:
 * co[0]=alpha, co[1..5]=beta1..5, co[6..8]=K1,K2,K3, co[9]=bt */

/* Pointwise Jacobian integrand of assemble.cc:229-252 for one (i, j) pair. */
static double lit_jac_point(const double *co, const double *u, const double *v, const double *gpi,
                            const double *gqi, const double *gpj, const double *gqj, const double *pi,
                            const double *pj, const double *qi, const double *qj)
{
  return ((co[6] * lit_mat_K1(gpi, gqi, gpj, gqj)) + ((co[7] + co[8]) * lit_mat_K2K3(gpi, gqi, gpj, gqj)) +
          (co[0] * lit_mat_alpha(pi, pj, qi, qj)) +
          2.0 * ((co[1] * lit_mat_b1(u, v, pi, pj, qi, qj)) + (co[2] * lit_mat_b2(u, v, pi, pj, qi, qj)) +
                 (co[3] * lit_mat_b3(u, v, pi, pj, qi, qj)) + (co[4] * lit_mat_b4(u, v, pi, pj, qi, qj)) +
                 (co[5] * lit_mat_b5(u, v, pi, pj, qi, qj))));
}

/* Pointwise residual integrand of assemble.cc:257-276 / residual.cc:206-226 */
static double lit_res_point(const double *co, const double *u, const double *v, const double *gu,
                            const double *gv, const double *pi, const double *qi, const double *gpi,
                            const double *gqi)
{
  return ((co[6] * lit_vec_K1(gu, gv, gpi, gqi)) + ((co[7] + co[8]) * lit_vec_K2K3(gpi, gqi, gu, gv)) +
          (co[0] * lit_vec_alpha(pi, qi, u, v)) +
          2.0 * ((co[1] * lit_vec_b1(u, v, pi, qi)) + (co[2] * lit_vec_b2(u, v, pi, qi)) +
                 (co[3] * lit_vec_b3(u, v, pi, qi)) + (co[4] * lit_vec_b4(u, v, pi, qi)) +
                 (co[5] * lit_vec_b5(u, v, pi, qi))));
}

/* unit tensor E_c split into (phi_u, phi_v) */
static void unit_phi(int c, double s, double *pu, double *pv)
{
  memset(pu, 0, 9 * sizeof(double));
  memset(pv, 0, 9 * sizeof(double));
  if (c < 9) pu[c] = s; else pv[c - 9] = s;
}

/* H_cd = alpha delta_cd + 2 sum_n beta_n m_n(u, v; E_c, E_d), evaluated
 * with the literal bilinear forms (bulk part of assemble.cc:234-252). */
void orc_point_hessian_literal(const double *co, const double *x, double *H)
{
  const double *u = x, *v = x + 9;
  double pi[9], qi[9], pj[9], qj[9];
  double g0[27];
  memset(g0, 0, sizeof g0);
  for (int c = 0; c < NC; ++c) {
    unit_phi(c, 1.0, pi, qi);
    for (int d = 0; d < NC; ++d) {
      unit_phi(d, 1.0, pj, qj);
      H[c * NC + d] = lit_jac_point(co, u, v, g0, g0, g0, g0, pi, pj, qi, qj) -
                      0.0; /* gradient terms vanish for zero gradients */
    }
  }
}

/* rho_c = alpha x_c + 2 sum_n beta_n r_n(u, v; E_c) (bulk part of the residual) */
void orc_point_rho_literal(const double *co, const double *x, double *rho)
{
  const double *u = x, *v = x + 9;
  double pi[9], qi[9], g0[27];
  memset(g0, 0, sizeof g0);
  for (int c = 0; c < NC; ++c) {
    unit_phi(c, 1.0, pi, qi);
    rho[c] = lit_res_point(co, u, v, g0, g0, pi, qi, g0, g0);
  }
}

/* literal pointwise integrands exposed for tests: J(i,j) and R(i) densities for
 * arbitrary 3x3 phi tensors (not only unit ones). */
double orc_lit_jac_point(const double *co, const double *x, const double *gphi_i, const double *gphi_j,
                         const double *phi_i, const double *phi_j)
{
  return lit_jac_point(co, x, x + 9, gphi_i, gphi_i + 27, gphi_j, gphi_j + 27, phi_i, phi_j, phi_i + 9,
                       phi_j + 9);
}
double orc_lit_res_point(const double *co, const double *x, const double *gx, const double *phi,
                         const double *gphi)
{
  return lit_res_point(co, x, x + 9, gx, gx + 27, phi, phi + 9, gphi, gphi + 27);
}

/* ------------------------------------------------------------------ */
/* Closed-form point Hessian and gradient of the bulk energy density.  */
/* f_bulk = alpha|x|^2 + sum_n beta_n I_n with I_1 = |Tr AA^T|^2,       */
/* I_2 = (Tr AA^+)^2, I_3 = Tr P^2 + Tr Q^2, I_4 = Tr S^2 - Tr W^2,     */
/* I_5 = Tr S^2 + Tr W^2 (P = uu^T - vv^T, Q = uv^T + vu^T,            */
/* S = uu^T + vv^T, W = vu^T - uv^T).  The literal forms above satisfy  */
/* r_n = (1/4) dI_n and m_n = (1/4) d^2 I_n (checked by the tests), so  */
/* rho = (1/2) grad f_bulk and H = (1/2) Hess f_bulk.                   */
/* ------------------------------------------------------------------ */
void orc_point_hessian(const double *co, const double *x, double *H)
{
  const double *u = x, *v = x + 9;
  double al = co[0], b1 = co[1], b2 = co[2], b3 = co[3], b4 = co[4], b5 = co[5];
  double U2 = trmT(u, u), V2 = trmT(v, v), UV = trmT(u, v);
  double D = U2 - V2, N = U2 + V2;
  double Auu[9], Avv[9], Auv[9], uut[9], vvt[9], uvt[9];
  Tm(u, u, Auu); Tm(v, v, Avv); Tm(u, v, Auv);   /* orbital Gram (k,l) */
  mT(u, u, uut); mT(v, v, vvt); mT(u, v, uvt);   /* spin Gram (mu,nu) */
  double b345p = b3 + b4 + b5;
  for (int mu = 0; mu < 3; ++mu)
    for (int k = 0; k < 3; ++k)
      for (int nu = 0; nu < 3; ++nu)
        for (int l = 0; l < 3; ++l) {
          int a = 3 * mu + k, b = 3 * nu + l;
          double dmn = (mu == nu), dkl = (k == l), dab = dmn * dkl;
          double ua = u[a], ub = u[b], va = v[a], vb = v[b];
          double uml = u[3 * mu + l], unk = u[3 * nu + k], vml = v[3 * mu + l], vnk = v[3 * nu + k];
          double Wmn = v[3 * mu + 0] * u[3 * nu + 0] + v[3 * mu + 1] * u[3 * nu + 1] +
                       v[3 * mu + 2] * u[3 * nu + 2] - uvt[3 * mu + nu]; /* (vu^T - uv^T)_{mu nu} */
          double Qmn = uvt[3 * mu + nu] + uvt[3 * nu + mu];
          /* uu block */
          double huu = dab * (al + 2 * b1 * D + 2 * b2 * N) + 4 * (b1 + b2) * ua * ub + 4 * b1 * va * vb +
                       2 * dmn * (b345p * Auu[3 * l + k] + (b3 + b4 - b5) * Avv[3 * l + k]) +
                       2 * b345p * uml * unk + 2 * (b3 - b4 + b5) * vml * vnk +
                       2 * dkl * (b345p * uut[3 * mu + nu] + (-b3 + b4 + b5) * vvt[3 * mu + nu]);
          /* vv block */
          double hvv = dab * (al - 2 * b1 * D + 2 * b2 * N) + 4 * (b1 + b2) * va * vb + 4 * b1 * ua * ub +
                       2 * dmn * (b345p * Avv[3 * l + k] + (b3 + b4 - b5) * Auu[3 * l + k]) +
                       2 * b345p * vml * vnk + 2 * (b3 - b4 + b5) * uml * unk +
                       2 * dkl * (b345p * vvt[3 * mu + nu] + (-b3 + b4 + b5) * uut[3 * mu + nu]);
          /* uv block: row index in u (mu,k), column index in v (nu,l) */
          double Auv_lk = Auv[3 * l + k], Avu_lk = Auv[3 * k + l]; /* (u^T v)_{lk}, (v^T u)_{lk} */
          double huv = dab * (2 * b1 * 2 * UV) + 4 * b1 * (ub * va - vb * ua) + 4 * b2 * ua * vb +
                       2 * b3 * (-dmn * Avu_lk - vml * unk + dmn * Auv_lk + uml * vnk + Qmn * dkl) +
                       2 * b4 * (dmn * Avu_lk + vml * unk - dmn * Auv_lk + uml * vnk - Wmn * dkl) +
                       2 * b5 * (dmn * Avu_lk + vml * unk + dmn * Auv_lk - uml * vnk + Wmn * dkl);
          H[a * NC + b] = huu;
          H[(a + 9) * NC + (b + 9)] = hvv;
          H[a * NC + (b + 9)] = huv;
          H[(b + 9) * NC + a] = huv;
        }
}

/* rho = alpha x + sum_n beta_n * (1/2) grad I_n (closed form) */
void orc_point_rho(const double *co, const double *x, double *rho)
{
  const double *u = x, *v = x + 9;
  double al = co[0], b1 = co[1], b2 = co[2], b3 = co[3], b4 = co[4], b5 = co[5];
  double U2 = trmT(u, u), V2 = trmT(v, v), UV = trmT(u, v), D = U2 - V2, N = U2 + V2;
  double uut[9], vvt[9], uvt[9], vut[9], P[9], Q[9], S[9], W[9];
  mT(u, u, uut); mT(v, v, vvt); mT(u, v, uvt); mT(v, u, vut);
  for (int k = 0; k < 9; ++k) {
    P[k] = uut[k] - vvt[k]; Q[k] = uvt[k] + vut[k];
    S[k] = uut[k] + vvt[k]; W[k] = vut[k] - uvt[k];
  }
  double Pu[9], Pv[9], Qu[9], Qv[9], Su[9], Sv[9], Wu[9], Wv[9];
  mm(P, u, Pu); mm(P, v, Pv); mm(Q, u, Qu); mm(Q, v, Qv);
  mm(S, u, Su); mm(S, v, Sv); mm(W, u, Wu); mm(W, v, Wv);
  for (int a = 0; a < 9; ++a) {
    /* G_n = (1/4) grad I_n */
    double g1u = D * u[a] + 2 * UV * v[a], g1v = -D * v[a] + 2 * UV * u[a];
    double g2u = N * u[a], g2v = N * v[a];
    double g3u = Pu[a] + Qv[a], g3v = Qu[a] - Pv[a];
    double g4u = Su[a] - Wv[a], g4v = Sv[a] + Wu[a];
    double g5u = Su[a] + Wv[a], g5v = Sv[a] - Wu[a];
    rho[a] = al * u[a] + 2 * (b1 * g1u + b2 * g2u + b3 * g3u + b4 * g4u + b5 * g5u);
    rho[a + 9] = al * v[a] + 2 * (b1 * g1v + b2 * g2v + b3 * g3v + b4 * g4v + b5 * g5v);
  }
}

/* bulk + gradient free-energy density at a point (free energy implied by
 * the residual: R = (1/2) dF/dU). gx[k*18 + c] = d_k x_c. */
double orc_point_energy(const double *co, const double *x, const double *gx)
{
  const double *u = x, *v = x + 9;
  double U2 = trmT(u, u), V2 = trmT(v, v), UV = trmT(u, v), D = U2 - V2, N = U2 + V2;
  double uut[9], vvt[9], uvt[9], vut[9], P[9], Q[9], S[9], W[9];
  mT(u, u, uut); mT(v, v, vvt); mT(u, v, uvt); mT(v, u, vut);
  for (int k = 0; k < 9; ++k) {
    P[k] = uut[k] - vvt[k]; Q[k] = uvt[k] + vut[k];
    S[k] = uut[k] + vvt[k]; W[k] = vut[k] - uvt[k];
  }
  double trP2 = trmm(P, P), trQ2 = trmm(Q, Q), trS2 = trmm(S, S), trW2 = trmm(W, W);
  double I1 = D * D + 4 * UV * UV, I2 = N * N, I3 = trP2 + trQ2, I4 = trS2 - trW2, I5 = trS2 + trW2;
  double f = co[0] * N + co[1] * I1 + co[2] * I2 + co[3] * I3 + co[4] * I4 + co[5] * I5;
  double g2 = 0;
  for (int k = 0; k < 3; ++k)
    for (int c = 0; c < NC; ++c) g2 += gx[k * NC + c] * gx[k * NC + c];
  double dv = 0;
  for (int part = 0; part < 2; ++part)
    for (int m = 0; m < 3; ++m) {
      double d = 0;
      for (int j = 0; j < 3; ++j) d += gx[j * NC + 9 * part + 3 * m + j];
      dv += d * d;
    }
  return f + co[6] * g2 + (co[7] + co[8]) * dv;
}

/* ------------------------------------------------------------------ */
/* Reference element: QGauss(k+1) and FE_Q(k) on axis-aligned boxes    */
/* (assemble.cc:123-132; QGauss<dim-1>(k+1) for faces, :124)          */
/* ------------------------------------------------------------------ */
static int gauss01(int n, double *x, double *w)
{
  if (n == 2) {
    double d = 0.5 / sqrt(3.0);
    x[0] = 0.5 - d; x[1] = 0.5 + d; w[0] = w[1] = 0.5;
  } else if (n == 3) {
    double d = 0.5 * sqrt(0.6);
    x[0] = 0.5 - d; x[1] = 0.5; x[2] = 0.5 + d;
    w[0] = w[2] = 5.0 / 18.0; w[1] = 8.0 / 18.0;
  } else
    return -1;
  return 0;
}
/* 1D Lagrange basis on equispaced support points (FE_Q(1), FE_Q(2) whose
 * Gauss-Lobatto points are 0, 1/2, 1) */
static void lag1d(int deg, double x, double *L, double *dL)
{
  if (deg == 1) {
    L[0] = 1 - x; L[1] = x; dL[0] = -1; dL[1] = 1;
  } else {
    L[0] = 2 * (x - 0.5) * (x - 1); L[1] = -4 * x * (x - 1); L[2] = 2 * x * (x - 0.5);
    dL[0] = 4 * x - 3; dL[1] = 4 - 8 * x; dL[2] = 4 * x - 1;
  }
}

typedef struct {
  int deg, n1, ns, nq;
  double N[27 * 27];       /* N[q*ns + a] */
  double dN[27 * 27 * 3];  /* dN[(q*ns + a)*3 + k] on the unit cube */
  double w[27];            /* reference weights */
} refel;

static int refel_init(refel *r, int deg)
{
  if (deg != 1 && deg != 2) return -1;
  r->deg = deg; r->n1 = deg + 1; r->ns = r->n1 * r->n1 * r->n1; r->nq = r->ns;
  double gx[3], gw[3], L[3][3], dL[3][3];
  gauss01(r->n1, gx, gw);
  for (int i = 0; i < r->n1; ++i) lag1d(deg, gx[i], L[i], dL[i]);
  int n1 = r->n1;
  for (int qz = 0; qz < n1; ++qz)
    for (int qy = 0; qy < n1; ++qy)
      for (int qx = 0; qx < n1; ++qx) {
        int q = qx + n1 * (qy + n1 * qz);
        r->w[q] = gw[qx] * gw[qy] * gw[qz];
        for (int az = 0; az < n1; ++az)
          for (int ay = 0; ay < n1; ++ay)
            for (int ax = 0; ax < n1; ++ax) {
              int a = ax + n1 * (ay + n1 * az);
              r->N[q * r->ns + a] = L[qx][ax] * L[qy][ay] * L[qz][az];
              r->dN[(q * r->ns + a) * 3 + 0] = dL[qx][ax] * L[qy][ay] * L[qz][az];
              r->dN[(q * r->ns + a) * 3 + 1] = L[qx][ax] * dL[qy][ay] * L[qz][az];
              r->dN[(q * r->ns + a) * 3 + 2] = L[qx][ax] * L[qy][ay] * dL[qz][az];
            }
      }
  return 0;
}

/* reference data for tests */
int orc_refel(int deg, double *N, double *dN, double *w)
{
  refel r;
  if (refel_init(&r, deg)) return -1;
  memcpy(N, r.N, sizeof(double) * r.nq * r.ns);
  memcpy(dN, r.dN, sizeof(double) * r.nq * r.ns * 3);
  memcpy(w, r.w, sizeof(double) * r.nq);
  return 0;
}

/* iterate and its gradient at quadrature point q (s_vector2matrix.cc:116-217) */
static void interp_q(const refel *r, const double *h, const double *Uloc, int q, double *x, double *gx)
{
  for (int c = 0; c < NC; ++c) {
    x[c] = 0; gx[c] = gx[NC + c] = gx[2 * NC + c] = 0;
  }
  for (int a = 0; a < r->ns; ++a) {
    double n = r->N[q * r->ns + a];
    double d0 = r->dN[(q * r->ns + a) * 3 + 0] / h[0];
    double d1 = r->dN[(q * r->ns + a) * 3 + 1] / h[1];
    double d2 = r->dN[(q * r->ns + a) * 3 + 2] / h[2];
    for (int c = 0; c < NC; ++c) {
      double U = Uloc[a * NC + c];
      x[c] += n * U; gx[c] += d0 * U; gx[NC + c] += d1 * U; gx[2 * NC + c] += d2 * U;
    }
  }
}

/* Faithful cell assembly: literal transliteration of assemble.cc:177-280.
 * Jloc[dpc*dpc] (row-major), Rloc[dpc] = cell_rhs = -R_cell.
 * Either output may be NULL.  Only for tiny meshes (cost ~dpc^2 n_q term
 * calls). */
int orc_cell_faithful(int deg, const double *h, const double *co, const double *Uloc, double *Jloc,
                      double *Rloc)
{
  refel r;
  if (refel_init(&r, deg)) return -1;
  int dpc = r.ns * NC;
  if (Jloc) memset(Jloc, 0, sizeof(double) * dpc * dpc);
  if (Rloc) memset(Rloc, 0, sizeof(double) * dpc);
  double vol = h[0] * h[1] * h[2];
  double x[NC], gx[3 * NC];
  double u[9], v[9], gu[27], gv[27];
  double pi[9], qi[9], pj[9], qj[9], gpi[27], gqi[27], gpj[27], gqj[27];
  for (int q = 0; q < r.nq; ++q) {
    double JxW = r.w[q] * vol;
    interp_q(&r, h, Uloc, q, x, gx);
    for (int c = 0; c < 9; ++c) { u[c] = x[c]; v[c] = x[9 + c]; }
    for (int k = 0; k < 3; ++k)
      for (int c = 0; c < 9; ++c) { gu[9 * k + c] = gx[k * NC + c]; gv[9 * k + c] = gx[k * NC + 9 + c]; }
    for (int i = 0; i < dpc; ++i) {
      int ai = i / NC, ci = i % NC;
      double ni = r.N[q * r.ns + ai];
      unit_phi(ci, ni, pi, qi);
      for (int k = 0; k < 3; ++k)
        unit_phi(ci, r.dN[(q * r.ns + ai) * 3 + k] / h[k], gpi + 9 * k, gqi + 9 * k);
      if (Jloc) {
        for (int j = 0; j < dpc; ++j) {
          int aj = j / NC, cj = j % NC;
          double nj = r.N[q * r.ns + aj];
          unit_phi(cj, nj, pj, qj);
          for (int k = 0; k < 3; ++k)
            unit_phi(cj, r.dN[(q * r.ns + aj) * 3 + k] / h[k], gpj + 9 * k, gqj + 9 * k);
          Jloc[(size_t)i * dpc + j] += lit_jac_point(co, u, v, gpi, gqi, gpj, gqj, pi, pj, qi, qj) * JxW;
        }
      }
      if (Rloc) Rloc[i] -= lit_res_point(co, u, v, gu, gv, pi, qi, gpi, gqi) * JxW;
    }
  }
  return 0;
}

/* Face geometry: local face f = 0:-x 1:+x 2:-y 3:+y 4:-z 5:+z. */
static void face_point(const refel *r, int f, int q2, double *xi, double *wf, const double *h)
{
  double gx[3], gw[3];
  gauss01(r->n1, gx, gw);
  int dn = f / 2, t0 = (dn + 1) % 3, t1 = (dn + 2) % 3;
  if (t0 > t1) { int t = t0; t0 = t1; t1 = t; }
  int i0 = q2 % r->n1, i1 = q2 / r->n1;
  xi[dn] = (f & 1) ? 1.0 : 0.0;
  xi[t0] = gx[i0];
  xi[t1] = gx[i1];
  *wf = gw[i0] * gw[i1] * h[t0] * h[t1];
}

/* normal-column mask of s_vector2matrix.cc:259-285 / phi_vector2matrix.cc:162-188:
 * boundary id 2 zeroes orbital column 0, id 3 column 1, id 4 column 2 */
static int face_zeroed(int bid, int c)
{
  int col = (c % 9) % 3;
  return (bid == 2 && col == 0) || (bid == 3 && col == 1) || (bid == 4 && col == 2);
}

/* Robin (AdGR diffuse) face term of assemble.cc:286-348 and residual.cc:235-278
 * for one boundary face; adds into Jloc/Rloc.  Active only when bt < 1e10 and
 * bid in {2,3,4}; no-op otherwise. */
int orc_face_faithful(int deg, const double *h, int f, int bid, const double *co, const double *Uloc,
                      double *Jloc, double *Rloc)
{
  double bt = co[9];
  if (!(bt < 1e10) || bid < 2 || bid > 4) return 0;
  refel r;
  if (refel_init(&r, deg)) return -1;
  int dpc = r.ns * NC, nqf = r.n1 * r.n1;
  double K1 = co[6];
  for (int q = 0; q < nqf; ++q) {
    double xi[3], wf, L[3][3], dL[3][3];
    face_point(&r, f, q, xi, &wf, h);
    for (int k = 0; k < 3; ++k) lag1d(deg, xi[k], L[k], dL[k]);
    double Ns[27];
    for (int az = 0; az < r.n1; ++az)
      for (int ay = 0; ay < r.n1; ++ay)
        for (int ax = 0; ax < r.n1; ++ax) Ns[ax + r.n1 * (ay + r.n1 * az)] = L[0][ax] * L[1][ay] * L[2][az];
    double uf[9], vf[9];
    for (int c = 0; c < NC; ++c) {
      double s = 0;
      if (!face_zeroed(bid, c))
        for (int a = 0; a < r.ns; ++a) s += Ns[a] * Uloc[a * NC + c];
      if (c < 9) uf[c] = s; else vf[c - 9] = s;
    }
    for (int i = 0; i < dpc; ++i) {
      int ai = i / NC, ci = i % NC;
      double pi[9], qi[9], pj[9], qj[9];
      unit_phi(ci, face_zeroed(bid, ci) ? 0.0 : Ns[ai], pi, qi);
      if (Jloc)
        for (int j = 0; j < dpc; ++j) {
          int aj = j / NC, cj = j % NC;
          unit_phi(cj, face_zeroed(bid, cj) ? 0.0 : Ns[aj], pj, qj);
          /* cell_mat_face_lhs_K1.cc:109-133, signs of assemble.cc:322-326 */
          Jloc[(size_t)i * dpc + j] -= (K1 * (-1.0 / bt) * (trmT(pi, pj) + trmT(qi, qj))) * wf;
        }
      /* cell_vec_face_rhs_K1.cc:109-133, signs of assemble.cc:333-337 */
      if (Rloc) Rloc[i] -= ((-K1) * (-1.0 / bt) * (trmT(uf, pi) + trmT(vf, qi))) * wf;
    }
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* Factorized cell assembly (same bilinear forms, one Hessian per q).  */
/* ------------------------------------------------------------------ */
int orc_cell_factorized(int deg, const double *h, const double *co, const double *Uloc, double *Jloc,
                        double *Rloc)
{
  refel r;
  if (refel_init(&r, deg)) return -1;
  int ns = r.ns, dpc = ns * NC;
  if (Jloc) memset(Jloc, 0, sizeof(double) * dpc * dpc);
  if (Rloc) memset(Rloc, 0, sizeof(double) * dpc);
  double vol = h[0] * h[1] * h[2];
  double K1 = co[6], K23 = co[7] + co[8];
  double x[NC], gx[3 * NC], H[NC * NC], rho[NC];
  for (int q = 0; q < r.nq; ++q) {
    double JxW = r.w[q] * vol;
    interp_q(&r, h, Uloc, q, x, gx);
    if (Jloc) orc_point_hessian(co, x, H);
    if (Rloc) orc_point_rho(co, x, rho);
    /* gamma_{c,k} = K1 d_k x_c + K23 [k == orbital(c)] d_part(mu) */
    double gam[3 * NC];
    for (int c = 0; c < NC; ++c) {
      int part = c / 9, mu = (c % 9) / 3, kc = c % 3;
      double dv = 0;
      for (int j = 0; j < 3; ++j) dv += gx[j * NC + 9 * part + 3 * mu + j];
      for (int k = 0; k < 3; ++k) gam[k * NC + c] = K1 * gx[k * NC + c] + (k == kc ? K23 * dv : 0.0);
    }
    for (int a = 0; a < ns; ++a) {
      double na = r.N[q * ns + a];
      double da[3];
      for (int k = 0; k < 3; ++k) da[k] = r.dN[(q * ns + a) * 3 + k] / h[k];
      if (Rloc)
        for (int c = 0; c < NC; ++c)
          Rloc[a * NC + c] -= JxW * (na * rho[c] + da[0] * gam[c] + da[1] * gam[NC + c] + da[2] * gam[2 * NC + c]);
      if (!Jloc) continue;
      for (int b = 0; b < ns; ++b) {
        double nb = r.N[q * ns + b];
        double db[3];
        for (int k = 0; k < 3; ++k) db[k] = r.dN[(q * ns + b) * 3 + k] / h[k];
        double wnn = JxW * na * nb;
        double lap = JxW * K1 * (da[0] * db[0] + da[1] * db[1] + da[2] * db[2]);
        for (int c = 0; c < NC; ++c) {
          double *row = Jloc + (size_t)(a * NC + c) * dpc + b * NC;
          for (int d = 0; d < NC; ++d) row[d] += wnn * H[c * NC + d];
          row[c] += lap;
          int part = c / 9, mu = (c % 9) / 3, kc = c % 3;
          for (int l = 0; l < 3; ++l) row[9 * part + 3 * mu + l] += JxW * K23 * da[kc] * db[l];
        }
      }
    }
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* Global assembly with affine constraints, as                          */
/* AffineConstraints::distribute_local_to_global (assemble.cc:356-361,  */
/* residual.cc:286-289): constrained rows/cols are distributed to their */
/* masters with the constraint weights; a constrained row receives the  */
/* |local diagonal| of each cell (or the local average |diagonal| when  */
/* that entry is zero), and a zero right-hand side.                    */
/* ------------------------------------------------------------------ */
typedef struct {
  int n_dofs;
  const int *con_of;      /* per dof: -1 unconstrained, else index into ptr */
  const int *ptr;         /* masters of constraint i: [ptr[i], ptr[i+1]) */
  const int *master;
  const double *weight;
} cons_t;

static int bsr_find(const int *row_ptr, const int *col, int r, int s)
{
  int lo = row_ptr[r], hi = row_ptr[r + 1] - 1;
  while (lo <= hi) {
    int mid = (lo + hi) >> 1;
    if (col[mid] == s) return mid;
    if (col[mid] < s) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

static int bsr_add(const int *row_ptr, const int *col, double *val, int gr, int gc, double v)
{
  int blk = bsr_find(row_ptr, col, gr / NC, gc / NC);
  if (blk < 0) return -1;
  val[(size_t)blk * NC * NC + (gr % NC) * NC + (gc % NC)] += v;
  return 0;
}

static void expand(const cons_t *cs, int g, int *m, double *w, int *n)
{
  int ci = cs->con_of[g];
  if (ci < 0) { m[0] = g; w[0] = 1.0; *n = 1; return; }
  *n = 0;
  for (int p = cs->ptr[ci]; p < cs->ptr[ci + 1]; ++p) { m[*n] = cs->master[p]; w[*n] = cs->weight[p]; ++*n; }
}

/*
 * Global Jacobian and right-hand side (system_rhs = -R) into a caller-owned
 * BSR(18x18) pattern over nodes.  Mesh:
 *   cell_nodes[n_cells*ns] (lexicographic local order), cell_h[n_cells*3],
 *   faces: n_faces entries (cell, local face 0..5, boundary id) for the
 *   Robin term, constraints as con_of/ptr/master/weight over dofs
 *   (dof = node*18 + c).
 * flags: bit0 = Jacobian, bit1 = rhs, bit2 = faithful cell form.
 * Returns 0, -1 bad degree, -2 an entry missing from the pattern.
 */
int orc_assemble(int deg, int n_cells, const int *cell_nodes, const double *cell_h, int n_faces,
                 const int *faces, int n_nodes, const int *con_of, const int *con_ptr, const int *con_master,
                 const double *con_weight, const double *co, const double *U, int flags, const int *row_ptr,
                 const int *col, double *val, double *rhs)
{
  refel r;
  if (refel_init(&r, deg)) return -1;
  int ns = r.ns, dpc = ns * NC;
  cons_t cs = {n_nodes * NC, con_of, con_ptr, con_master, con_weight};
  int doJ = flags & 1, doR = (flags >> 1) & 1, faithful = (flags >> 2) & 1;
  if (doJ) memset(val, 0, sizeof(double) * (size_t)row_ptr[n_nodes] * NC * NC);
  if (doR) memset(rhs, 0, sizeof(double) * (size_t)n_nodes * NC);
  double *Jl = doJ ? (double *)malloc(sizeof(double) * dpc * dpc) : NULL;
  double *Rl = (double *)malloc(sizeof(double) * dpc);
  double *Ul = (double *)malloc(sizeof(double) * dpc);
  int *gdof = (int *)malloc(sizeof(int) * dpc);
  int rc = 0;
  for (int e = 0; e < n_cells && !rc; ++e) {
    const int *cn = cell_nodes + (size_t)e * ns;
    const double *h = cell_h + 3 * e;
    for (int a = 0; a < ns; ++a)
      for (int c = 0; c < NC; ++c) {
        gdof[a * NC + c] = cn[a] * NC + c;
        Ul[a * NC + c] = U[(size_t)cn[a] * NC + c];
      }
    if (faithful) orc_cell_faithful(deg, h, co, Ul, Jl, Rl);
    else orc_cell_factorized(deg, h, co, Ul, Jl, Rl);
    for (int f = 0; f < n_faces; ++f)
      if (faces[3 * f] == e) orc_face_faithful(deg, h, faces[3 * f + 1], faces[3 * f + 2], co, Ul, Jl, Rl);
    /* scatter */
    double avg = 0;
    if (doJ) {
      for (int i = 0; i < dpc; ++i) avg += fabs(Jl[(size_t)i * dpc + i]);
      avg /= dpc;
    }
    for (int i = 0; i < dpc && !rc; ++i) {
      int mi[64], ni; double wi[64];
      expand(&cs, gdof[i], mi, wi, &ni);
      if (doR)
        for (int p = 0; p < ni; ++p) rhs[mi[p]] += wi[p] * Rl[i];
      if (!doJ) continue;
      for (int j = 0; j < dpc && !rc; ++j) {
        int mj[64], nj; double wj[64];
        expand(&cs, gdof[j], mj, wj, &nj);
        double v = Jl[(size_t)i * dpc + j];
        for (int p = 0; p < ni && !rc; ++p)
          for (int s = 0; s < nj && !rc; ++s)
            if (bsr_add(row_ptr, col, val, mi[p], mj[s], wi[p] * wj[s] * v)) rc = -2;
      }
      if (con_of[gdof[i]] >= 0) {
        double d = fabs(Jl[(size_t)i * dpc + i]);
        if (d == 0) d = avg;
        if (bsr_add(row_ptr, col, val, gdof[i], gdof[i], d)) rc = -2;
      }
    }
  }
  free(Jl); free(Rl); free(Ul); free(gdof);
  return rc;
}

/*
 * Matrix-free check at scale: system_rhs = -R and Y = J X for nvec vectors
 * (X, Y: nvec * n_nodes*18, vector-major) without forming the global J,
 * cell by cell with the factorized cell form and the Robin faces, scattered
 * as distribute_local_to_global does for homogeneous Dirichlet constraints
 * only (node_mask bit c = dof node*18 + c constrained; meshes without
 * hanging-node or periodic lines): constrained rows and columns drop out of
 * the cell contributions, a constrained row receives the sum over its cells
 * of |local diagonal| (local average when zero) times x, and a zero rhs
 * (assemble.cc:356-361, residual.cc:286-289).  Cells of one colour share no
 * node, so each colour runs in parallel over its cells (OpenMP) and the sums
 * are in a fixed order.  rhs or Y may be NULL.  Returns 0, or -1 on a bad
 * degree.
 */
int orc_cells_apply(int deg, int n_cells, const int *cell_nodes, const double *cell_h, int n_faces,
                    const int *faces, int n_nodes, const uint32_t *node_mask, const double *co, const double *U,
                    int n_colors, const int *color_ptr, const int *color_cells, int nvec, const double *X,
                    double *Y, double *rhs)
{
  refel r0;
  if (refel_init(&r0, deg)) return -1;
  const int ns = r0.ns, dpc = ns * NC;
  const size_t nd = (size_t)n_nodes * NC;
  const int doJ = Y && nvec > 0;
  if (rhs) memset(rhs, 0, sizeof(double) * nd);
  if (doJ) memset(Y, 0, sizeof(double) * nd * nvec);
  /* faces of each cell */
  int *fptr = (int *)calloc((size_t)n_cells + 1, sizeof(int)), *fidx = (int *)malloc(sizeof(int) * (n_faces + 1));
  for (int f = 0; f < n_faces; ++f) fptr[faces[3 * f] + 1]++;
  for (int e = 0; e < n_cells; ++e) fptr[e + 1] += fptr[e];
  {
    int *pos = (int *)malloc(sizeof(int) * ((size_t)n_cells + 1));
    memcpy(pos, fptr, sizeof(int) * ((size_t)n_cells + 1));
    for (int f = 0; f < n_faces; ++f) fidx[pos[faces[3 * f]]++] = f;
    free(pos);
  }
  for (int k = 0; k < n_colors; ++k) {
#ifdef _OPENMP
#pragma omp parallel
#endif
    {
      double *Jl = doJ ? (double *)malloc(sizeof(double) * dpc * dpc) : NULL;
      double Rl[27 * NC], Ul[27 * NC];
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 4)
#endif
      for (int ii = color_ptr[k]; ii < color_ptr[k + 1]; ++ii) {
        const int e = color_cells[ii];
        const int *cn = cell_nodes + (size_t)e * ns;
        const double *h = cell_h + 3 * e;
        for (int a = 0; a < ns; ++a)
          for (int c = 0; c < NC; ++c) Ul[a * NC + c] = U[(size_t)cn[a] * NC + c];
        orc_cell_factorized(deg, h, co, Ul, Jl, Rl);
        for (int f = fptr[e]; f < fptr[e + 1]; ++f)
          orc_face_faithful(deg, h, faces[3 * fidx[f] + 1], faces[3 * fidx[f] + 2], co, Ul, Jl, Rl);
        double avg = 0;
        if (doJ) {
          for (int i = 0; i < dpc; ++i) avg += fabs(Jl[(size_t)i * dpc + i]);
          avg /= dpc;
        }
        for (int i = 0; i < dpc; ++i) {
          const int ni = cn[i / NC], ci = i % NC;
          const size_t gi = (size_t)ni * NC + ci;
          const int con_i = (node_mask[ni] >> ci) & 1u;
          if (rhs && !con_i) rhs[gi] += Rl[i];
          if (!doJ) continue;
          if (con_i) {
            double d = fabs(Jl[(size_t)i * dpc + i]);
            if (d == 0) d = avg;
            for (int v = 0; v < nvec; ++v) Y[v * nd + gi] += d * X[v * nd + gi];
            continue;
          }
          for (int v = 0; v < nvec; ++v) {
            double s = 0;
            for (int j = 0; j < dpc; ++j) {
              const int nj = cn[j / NC], cj = j % NC;
              if ((node_mask[nj] >> cj) & 1u) continue;
              s += Jl[(size_t)i * dpc + j] * X[v * nd + (size_t)nj * NC + cj];
            }
            Y[v * nd + gi] += s;
          }
        }
      }
      free(Jl);
    }
  }
  free(fptr);
  free(fidx);
  return 0;
}

/* Free energy F = sum_cells sum_q JxW f(q) + Robin faces (K1/bt)|x~|^2. */
double orc_free_energy(int deg, int n_cells, const int *cell_nodes, const double *cell_h, int n_faces,
                       const int *faces, const double *co, const double *U)
{
  refel r;
  if (refel_init(&r, deg)) return NAN;
  int ns = r.ns;
  double F = 0, Ul[27 * NC], x[NC], gx[3 * NC];
  for (int e = 0; e < n_cells; ++e) {
    const int *cn = cell_nodes + (size_t)e * ns;
    const double *h = cell_h + 3 * e;
    double vol = h[0] * h[1] * h[2];
    for (int a = 0; a < ns; ++a)
      for (int c = 0; c < NC; ++c) Ul[a * NC + c] = U[(size_t)cn[a] * NC + c];
    double Fe = 0;
    for (int q = 0; q < r.nq; ++q) {
      interp_q(&r, h, Ul, q, x, gx);
      Fe += r.w[q] * vol * orc_point_energy(co, x, gx);
    }
    F += Fe;
  }
  double bt = co[9];
  if (bt < 1e10)
    for (int f = 0; f < n_faces; ++f) {
      int e = faces[3 * f], lf = faces[3 * f + 1], bid = faces[3 * f + 2];
      if (bid < 2 || bid > 4) continue;
      const int *cn = cell_nodes + (size_t)e * ns;
      const double *h = cell_h + 3 * e;
      for (int a = 0; a < ns; ++a)
        for (int c = 0; c < NC; ++c) Ul[a * NC + c] = U[(size_t)cn[a] * NC + c];
      for (int q = 0; q < r.n1 * r.n1; ++q) {
        double xi[3], wf, L[3][3], dL[3][3], s2 = 0;
        face_point(&r, lf, q, xi, &wf, h);
        for (int k = 0; k < 3; ++k) lag1d(deg, xi[k], L[k], dL[k]);
        for (int c = 0; c < NC; ++c) {
          if (face_zeroed(bid, c)) continue;
          double s = 0;
          for (int az = 0; az < r.n1; ++az)
            for (int ay = 0; ay < r.n1; ++ay)
              for (int ax = 0; ax < r.n1; ++ax)
                s += L[0][ax] * L[1][ay] * L[2][az] * Ul[(ax + r.n1 * (ay + r.n1 * az)) * NC + c];
          s2 += s * s;
        }
        F += wf * (co[6] / bt) * s2;
      }
    }
  return F;
}

/* ------------------------------------------------------------------ */
/* Linear algebra on BSR(18x18): SpMV, point-block Jacobi, GMRES(m).   */
/* The reference uses FGMRES + ML AMG (solve.cc:130-174); the accelerated */
/* path's solver is GMRES with 18x18 point-block Jacobi, restated here  */
/* as the CPU baseline of that solver.                                  */
/* ------------------------------------------------------------------ */
void orc_bsr_spmv(int n_nodes, const int *row_ptr, const int *col, const double *val, const double *x, double *y)
{
  for (int r = 0; r < n_nodes; ++r) {
    double acc[NC] = {0};
    for (int b = row_ptr[r]; b < row_ptr[r + 1]; ++b) {
      const double *B = val + (size_t)b * NC * NC;
      const double *xc = x + (size_t)col[b] * NC;
      for (int c = 0; c < NC; ++c) {
        double s = 0;
        for (int d = 0; d < NC; ++d) s += B[c * NC + d] * xc[d];
        acc[c] += s;
      }
    }
    memcpy(y + (size_t)r * NC, acc, sizeof acc);
  }
}

/* invert an 18x18 block (Gauss-Jordan with partial pivoting); -1 if singular */
static int inv18(const double *A, double *Ai)
{
  double M[NC][2 * NC];
  for (int i = 0; i < NC; ++i)
    for (int j = 0; j < NC; ++j) { M[i][j] = A[i * NC + j]; M[i][NC + j] = (i == j); }
  for (int c = 0; c < NC; ++c) {
    int p = c;
    for (int i = c + 1; i < NC; ++i) if (fabs(M[i][c]) > fabs(M[p][c])) p = i;
    if (M[p][c] == 0) return -1;
    if (p != c) for (int j = 0; j < 2 * NC; ++j) { double t = M[c][j]; M[c][j] = M[p][j]; M[p][j] = t; }
    double s = 1.0 / M[c][c];
    for (int j = 0; j < 2 * NC; ++j) M[c][j] *= s;
    for (int i = 0; i < NC; ++i)
      if (i != c && M[i][c] != 0) {
        double f = M[i][c];
        for (int j = 0; j < 2 * NC; ++j) M[i][j] -= f * M[c][j];
      }
  }
  for (int i = 0; i < NC; ++i) for (int j = 0; j < NC; ++j) Ai[i * NC + j] = M[i][NC + j];
  return 0;
}

int orc_block_jacobi(int n_nodes, const int *row_ptr, const int *col, const double *val, double *Dinv)
{
  for (int r = 0; r < n_nodes; ++r) {
    int b = bsr_find(row_ptr, col, r, r);
    if (b < 0) return -2;
    if (inv18(val + (size_t)b * NC * NC, Dinv + (size_t)r * NC * NC)) return -1;
  }
  return 0;
}

static void apply_bj(int n_nodes, const double *Dinv, const double *x, double *y)
{
  for (int r = 0; r < n_nodes; ++r)
    for (int c = 0; c < NC; ++c) {
      double s = 0;
      for (int d = 0; d < NC; ++d) s += Dinv[(size_t)r * NC * NC + c * NC + d] * x[(size_t)r * NC + d];
      y[(size_t)r * NC + c] = s;
    }
}

static double dot(size_t n, const double *a, const double *b)
{
  double s = 0;
  for (size_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

/* Right-preconditioned restarted GMRES(m) with point-block Jacobi and
 * classical Gram-Schmidt with one reorthogonalisation (CGS2).  Solves
 * J x = b from x = 0 until ||b - J x|| <= rtol ||b|| or maxit iterations.
 * Returns the iteration count (>= 0), or -1 when maxit is reached
 * (SolverControl::NoConvergence in the reference, solve.cc:159-174). */
int orc_gmres(int n_nodes, const int *row_ptr, const int *col, const double *val, const double *b, double *x,
              double rtol, int maxit, int m, double *res_out)
{
  size_t n = (size_t)n_nodes * NC;
  double *Dinv = (double *)malloc(sizeof(double) * (size_t)n_nodes * NC * NC);
  if (orc_block_jacobi(n_nodes, row_ptr, col, val, Dinv)) { free(Dinv); return -2; }
  double *V = (double *)malloc(sizeof(double) * n * (m + 1));
  double *Z = (double *)malloc(sizeof(double) * n);
  double *w = (double *)malloc(sizeof(double) * n);
  double *Hh = (double *)calloc((size_t)(m + 1) * m, sizeof(double));
  double *cs = (double *)malloc(sizeof(double) * m), *sn = (double *)malloc(sizeof(double) * m);
  double *g = (double *)malloc(sizeof(double) * (m + 1)), *yv = (double *)malloc(sizeof(double) * m);
  double *hc = (double *)malloc(sizeof(double) * (m + 1));
  memset(x, 0, sizeof(double) * n);
  double bn = sqrt(dot(n, b, b)), target = rtol * bn;
  int it = 0, conv = 0;
  double res = bn;
  while (it < maxit || it == 0) {
    /* true residual r = b - J x at every restart */
    orc_bsr_spmv(n_nodes, row_ptr, col, val, x, w);
    for (size_t i = 0; i < n; ++i) V[i] = b[i] - w[i];
    double beta = sqrt(dot(n, V, V));
    res = beta;
    if (beta <= target) { conv = 1; break; }
    if (it >= maxit) break;
    for (size_t i = 0; i < n; ++i) V[i] /= beta;
    memset(g, 0, sizeof(double) * (m + 1));
    g[0] = beta;
    int j = 0, est = 0;
    while (j < m && it < maxit && !est) {
      apply_bj(n_nodes, Dinv, V + n * j, Z);
      orc_bsr_spmv(n_nodes, row_ptr, col, val, Z, w);
      for (int i = 0; i <= j; ++i) Hh[i * m + j] = 0.0;
      for (int pass = 0; pass < 2; ++pass) {
        for (int i = 0; i <= j; ++i) hc[i] = dot(n, V + n * i, w);
        for (int i = 0; i <= j; ++i) {
          const double *vi = V + n * i;
          for (size_t k = 0; k < n; ++k) w[k] -= hc[i] * vi[k];
          Hh[i * m + j] += hc[i];
        }
      }
      double hn = sqrt(dot(n, w, w));
      if (hn > 0) for (size_t k = 0; k < n; ++k) V[n * (j + 1) + k] = w[k] / hn;
      for (int i = 0; i < j; ++i) {
        double a = Hh[i * m + j], c2 = Hh[(i + 1) * m + j];
        Hh[i * m + j] = cs[i] * a + sn[i] * c2;
        Hh[(i + 1) * m + j] = -sn[i] * a + cs[i] * c2;
      }
      double a = Hh[j * m + j], rr = hypot(a, hn);
      cs[j] = a / rr; sn[j] = hn / rr;
      Hh[j * m + j] = rr;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      ++j; ++it;
      if (fabs(g[j]) <= target || hn == 0) est = 1;
    }
    /* x += M^-1 V y, y from the triangular system */
    for (int i = j - 1; i >= 0; --i) {
      double s = g[i];
      for (int k = i + 1; k < j; ++k) s -= Hh[i * m + k] * yv[k];
      yv[i] = s / Hh[i * m + i];
    }
    memset(w, 0, sizeof(double) * n);
    for (int i = 0; i < j; ++i) for (size_t k = 0; k < n; ++k) w[k] += yv[i] * V[n * i + k];
    apply_bj(n_nodes, Dinv, w, Z);
    for (size_t k = 0; k < n; ++k) x[k] += Z[k];
  }
  if (res_out) *res_out = res;
  free(Dinv); free(V); free(Z); free(w); free(Hh); free(cs); free(sn); free(g); free(yv); free(hc);
  return conv ? it : -1;
}

/* ------------------------------------------------------------------ */
/* Constraint distribution (AffineConstraints::distribute, used at      */
/* iteration.cc:180 and solve.cc:181): constrained dofs take the        */
/* weighted sum of their masters (homogeneous: Dirichlet dofs -> 0).    */
/* ------------------------------------------------------------------ */
void orc_distribute(int n_dofs, const int *con_of, const int *con_ptr, const int *con_master,
                    const double *con_weight, double *x)
{
  for (int g = 0; g < n_dofs; ++g) {
    int ci = con_of[g];
    if (ci < 0) continue;
    double s = 0;
    for (int p = con_ptr[ci]; p < con_ptr[ci + 1]; ++p) s += con_weight[p] * x[con_master[p]];
    x[g] = s;
  }
}

/* ------------------------------------------------------------------ */
/* Seeded initial guess.  The reference draws from std::random_device   */
/* (setup_uniform_B-phase.cc:218-219), which is not reproducible; the   */
/* seeded form is: bulk B phase gap_B * 0.577350269f on u_11,u_22,u_33  */
/* (setup_uniform_B-phase.cc:245-259) plus amp * gap_B * (2 U - 1) on   */
/* every dof in node-major dof order, U = (mt19937_64() >> 11) * 2^-53. */
/* mt19937_64 is the published 64-bit Mersenne Twister (Matsumoto and  */
/* Nishimura), the sequence of C++11 std::mt19937_64.                   */
/* ------------------------------------------------------------------ */
typedef struct { uint64_t mt[312]; int i; } mt64;
static void mt64_seed(mt64 *s, uint64_t seed)
{
  s->mt[0] = seed;
  for (int i = 1; i < 312; ++i)
    s->mt[i] = 6364136223846793005ULL * (s->mt[i - 1] ^ (s->mt[i - 1] >> 62)) + (uint64_t)i;
  s->i = 312;
}
static uint64_t mt64_next(mt64 *s)
{
  static const uint64_t UM = 0xFFFFFFFF80000000ULL, LM = 0x7FFFFFFFULL, MA = 0xB5026F5AA96619E9ULL;
  if (s->i >= 312) {
    for (int k = 0; k < 312; ++k) {
      uint64_t y = (s->mt[k] & UM) | (s->mt[(k + 1) % 312] & LM);
      s->mt[k] = s->mt[(k + 156) % 312] ^ (y >> 1) ^ ((y & 1ULL) ? MA : 0ULL);
    }
    s->i = 0;
  }
  uint64_t y = s->mt[s->i++];
  y ^= (y >> 29) & 0x5555555555555555ULL;
  y ^= (y << 17) & 0x71D67FFFEDA60000ULL;
  y ^= (y << 37) & 0xFFF7EEE000000000ULL;
  y ^= (y >> 43);
  return y;
}

void orc_mt64(uint64_t seed, int n, uint64_t *out)
{
  mt64 s;
  mt64_seed(&s, seed);
  for (int i = 0; i < n; ++i) out[i] = mt64_next(&s);
}

/* phase: 0 = B phase (u_11 = u_22 = u_33 = gap * 0.577350269f), 1 = A phase
 * (u_11 = v_12 = gap * 0.707107f, setup_uniform_A-phase.cc:238-250) */
void orc_initial_guess(int n_nodes, double gap, int phase, double amp, uint64_t seed, double *U)
{
  mt64 s;
  mt64_seed(&s, seed);
  for (int nd = 0; nd < n_nodes; ++nd)
    for (int c = 0; c < NC; ++c) {
      double base = 0;
      if (phase == 0 && (c == 0 || c == 4 || c == 8)) base = gap * (double)0.577350269f;
      if (phase == 1 && (c == 0 || c == 10)) base = gap * (double)0.707107f;
      double r = (double)(mt64_next(&s) >> 11) * (1.0 / 9007199254740992.0);
      U[(size_t)nd * NC + c] = base + amp * gap * (2.0 * r - 1.0);
    }
}

/* A-phase / B-phase flat wall (femgl/inc/BnA.h:130-163): B phase at
   z >= z_a (u11 = u22 = u33 = gap_B * 0.5773502691896258f), A phase below
   (u11 = v12 = gap_A * 0.7071067811865475f), plus amp*gap_B*(2u-1) noise. */
void orc_initial_guess_bna(int n_nodes, const double *xyz, double z_a, double gap_A, double gap_B, double amp,
                           uint64_t seed, double *U)
{
  mt64 s;
  mt64_seed(&s, seed);
  for (int nd = 0; nd < n_nodes; ++nd) {
    int b = xyz[3 * nd + 2] >= z_a;
    for (int c = 0; c < NC; ++c) {
      double base = 0;
      if (b && (c == 0 || c == 4 || c == 8)) base = gap_B * (double)0.5773502691896258f;
      if (!b && (c == 0 || c == 10)) base = gap_A * (double)0.7071067811865475f;
      double r = (double)(mt64_next(&s) >> 11) * (1.0 / 9007199254740992.0);
      U[(size_t)nd * NC + c] = base + amp * gap_B * (2.0 * r - 1.0);
    }
  }
}

/* ------------------------------------------------------------------ */
/* Newton driver: assemble -> GMRES -> damped line search, as           */
/* run.cc:208-218 + iteration.cc:109-215, with the relative stopping    */
/* test ||R||/||R0|| <= rtol.                                           */
/* stats: [0] newton iterations, [1] total GMRES iterations,            */
/*        [2] ||R0||, [3] final ||R||.                                  */
/* ------------------------------------------------------------------ */
int orc_newton(int deg, int n_cells, const int *cell_nodes, const double *cell_h, int n_faces, const int *faces,
               int n_nodes, const int *con_of, const int *con_ptr, const int *con_master, const double *con_weight,
               const double *co, double *U, const int *row_ptr, const int *col, double rtol, int max_newton,
               double lin_tol, int lin_maxit, int restart, double ls_step, int damped, double *stats)
{
  size_t n = (size_t)n_nodes * NC, nb = (size_t)row_ptr[n_nodes];
  double *val = (double *)malloc(sizeof(double) * nb * NC * NC);
  double *rhs = (double *)malloc(sizeof(double) * n), *dU = (double *)malloc(sizeof(double) * n);
  double *xt = (double *)malloc(sizeof(double) * n), *rt = (double *)malloc(sizeof(double) * n);
  int rc = 0, k = 0, lin_total = 0;
  orc_distribute((int)n, con_of, con_ptr, con_master, con_weight, U);
  rc = orc_assemble(deg, n_cells, cell_nodes, cell_h, n_faces, faces, n_nodes, con_of, con_ptr, con_master,
                    con_weight, co, U, 2, row_ptr, col, val, rhs);
  double r0 = sqrt(dot(n, rhs, rhs)), rcur = r0;
  for (k = 0; k < max_newton && !rc; ++k) {
    if (rcur <= rtol * r0) break;
    rc = orc_assemble(deg, n_cells, cell_nodes, cell_h, n_faces, faces, n_nodes, con_of, con_ptr, con_master,
                      con_weight, co, U, 3, row_ptr, col, val, rhs);
    if (rc) break;
    double prev = sqrt(dot(n, rhs, rhs));
    double lres;
    int its = orc_gmres(n_nodes, row_ptr, col, val, rhs, dU, lin_tol, lin_maxit, restart, &lres);
    if (its < 0) { rc = -3; break; }
    lin_total += its;
    orc_distribute((int)n, con_of, con_ptr, con_master, con_weight, dU);
    int n_try = damped ? 100 : 1;
    for (int i = 0; i < n_try; ++i) {
      double a = damped ? pow(ls_step, (double)i) : 1.0;
      for (size_t t = 0; t < n; ++t) xt[t] = U[t] + a * dU[t];
      orc_distribute((int)n, con_of, con_ptr, con_master, con_weight, xt);
      orc_assemble(deg, n_cells, cell_nodes, cell_h, n_faces, faces, n_nodes, con_of, con_ptr, con_master,
                   con_weight, co, xt, 2, row_ptr, col, val, rt);
      rcur = sqrt(dot(n, rt, rt));
      if (rcur < prev) break;
    }
    memcpy(U, xt, sizeof(double) * n);
  }
  if (stats) { stats[0] = k; stats[1] = lin_total; stats[2] = r0; stats[3] = rcur; }
  free(val); free(rhs); free(dU); free(xt); free(rt);
  return rc;
}

/* ------------------------------------------------------------------ */
/* CPU-baseline timers (bench.py cpu_baseline leg): the oracle's cell  */
/* assembly and GMRES kernels over a bounded sample, OpenMP over cells */
/* and rows on the host cores.                                         */
/* ------------------------------------------------------------------ */
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif
static double wall(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* Local J and R of n sample cells (Uloc_all[n*dpc], h[3] shared), in
 * parallel over cells.  Returns wall seconds; checksum[0] receives a sum of
 * the results so the work cannot be elided. */
double orc_time_cells(int deg, int n, const double *h, const double *co, const double *Uloc_all, int faithful,
                      int nthreads, double *checksum)
{
  int ns = (deg + 1) * (deg + 1) * (deg + 1), dpc = ns * NC;
  double t0 = wall(), sum = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel reduction(+ : sum)
#endif
  {
    double *J = (double *)malloc(sizeof(double) * dpc * dpc), *R = (double *)malloc(sizeof(double) * dpc);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
    for (int e = 0; e < n; ++e) {
      if (faithful) orc_cell_faithful(deg, h, co, Uloc_all + (size_t)e * dpc, J, R);
      else orc_cell_factorized(deg, h, co, Uloc_all + (size_t)e * dpc, J, R);
      sum += J[(size_t)(e % dpc) * dpc] + R[e % dpc];
    }
    free(J); free(R);
  }
  if (checksum) checksum[0] = sum;
  return wall() - t0;
}

/* reps BSR SpMVs y = A x, parallel over block rows; returns wall seconds */
double orc_time_spmv(int n_nodes, const int *row_ptr, const int *col, const double *val, const double *x, double *y,
                     int reps, int nthreads)
{
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  double t0 = wall();
  for (int it = 0; it < reps; ++it) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int r = 0; r < n_nodes; ++r) {
      double acc[NC] = {0};
      for (int b = row_ptr[r]; b < row_ptr[r + 1]; ++b) {
        const double *B = val + (size_t)b * NC * NC, *xc = x + (size_t)col[b] * NC;
        for (int c = 0; c < NC; ++c) {
          double s = 0;
          for (int d = 0; d < NC; ++d) s += B[c * NC + d] * xc[d];
          acc[c] += s;
        }
      }
      memcpy(y + (size_t)r * NC, acc, sizeof acc);
    }
  }
  return wall() - t0;
}

/* Wall seconds of n_steps Arnoldi steps of the oracle's GMRES (solve.cc:
 * 159-174 with point-block Jacobi): z = D^-1 v_j, w = J z, one classical
 * Gram-Schmidt pass against v_0..v_j (the dots, then the update), the norm
 * and the scaling into v_{j+1}, from a seeded start vector; every loop over
 * rows, nodes or entries in parallel (OpenMP, nthreads).  The block-Jacobi
 * inversion is outside the timing.  Returns -1 on a singular block. */
double orc_time_arnoldi(int n_nodes, const int *row_ptr, const int *col, const double *val, int n_steps,
                        int nthreads)
{
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  const size_t n = (size_t)n_nodes * NC;
  double *Dinv = (double *)malloc(sizeof(double) * (size_t)n_nodes * NC * NC);
  if (orc_block_jacobi(n_nodes, row_ptr, col, val, Dinv)) { free(Dinv); return -1.0; }
  double *V = (double *)malloc(sizeof(double) * n * (n_steps + 1));
  double *Z = (double *)malloc(sizeof(double) * n), *w = (double *)malloc(sizeof(double) * n);
  double *h = (double *)malloc(sizeof(double) * (n_steps + 1));
  uint64_t st = 88172645463325252ull;
  for (size_t i = 0; i < n; ++i) {
    st ^= st << 13; st ^= st >> 7; st ^= st << 17;
    V[i] = (double)(st >> 11) * 0x1.0p-53 - 0.5;
  }
  double t0 = wall();
  for (int j = 0; j < n_steps; ++j) {
    const double *vj = V + n * j;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int r = 0; r < n_nodes; ++r)
      for (int c = 0; c < NC; ++c) {
        double s = 0;
        for (int d = 0; d < NC; ++d) s += Dinv[(size_t)r * NC * NC + c * NC + d] * vj[(size_t)r * NC + d];
        Z[(size_t)r * NC + c] = s;
      }
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int r = 0; r < n_nodes; ++r) {
      double acc[NC] = {0};
      for (int b = row_ptr[r]; b < row_ptr[r + 1]; ++b) {
        const double *B = val + (size_t)b * NC * NC, *zc = Z + (size_t)col[b] * NC;
        for (int c = 0; c < NC; ++c) {
          double s = 0;
          for (int d = 0; d < NC; ++d) s += B[c * NC + d] * zc[d];
          acc[c] += s;
        }
      }
      memcpy(w + (size_t)r * NC, acc, sizeof acc);
    }
    for (int i = 0; i <= j; ++i) {
      double s = 0;
      const double *vi = V + n * i;
#ifdef _OPENMP
#pragma omp parallel for reduction(+ : s) schedule(static)
#endif
      for (size_t k = 0; k < n; ++k) s += vi[k] * w[k];
      h[i] = s;
    }
    double nn = 0;
#ifdef _OPENMP
#pragma omp parallel for reduction(+ : nn) schedule(static)
#endif
    for (size_t k = 0; k < n; ++k) {
      double v = w[k];
      for (int i = 0; i <= j; ++i) v -= h[i] * V[n * i + k];
      w[k] = v;
      nn += v * v;
    }
    const double inv = nn > 0 ? 1.0 / sqrt(nn) : 0.0;
    double *vn = V + n * (j + 1);
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (size_t k = 0; k < n; ++k) vn[k] = w[k] * inv;
  }
  double el = wall() - t0;
  free(Dinv); free(V); free(Z); free(w); free(h);
  return el;
}

/* Residual-only cell work (compute_residual, residual.cc:166-229) of n
 * sample cells with the factorized form, in parallel; wall seconds. */
double orc_time_residual_cells(int deg, int n, const double *h, const double *co, const double *Uloc_all,
                               int nthreads, double *checksum)
{
  int ns = (deg + 1) * (deg + 1) * (deg + 1), dpc = ns * NC;
  double t0 = wall(), sum = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel reduction(+ : sum)
#endif
  {
    double *R = (double *)malloc(sizeof(double) * dpc);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 4)
#endif
    for (int e = 0; e < n; ++e) {
      orc_cell_factorized(deg, h, co, Uloc_all + (size_t)e * dpc, NULL, R);
      sum += R[e % dpc];
    }
    free(R);
  }
  if (checksum) checksum[0] = sum;
  return wall() - t0;
}

int orc_max_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
