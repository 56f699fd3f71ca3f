// Cell-loop assembly of the Ginzburg-Landau Jacobian J, its matrix-free form,
// and system_rhs = -R for FE_Q(k) x 18 on axis-aligned hexes, hand-written
// for gfx950.
//
// Reference: femgl/src/assemble.cc:177-280 (cell loop), residual.cc:152-232,
// the term functions of femgl/src/cell_mat_vec/, and the Robin face loops
// assemble.cc:286-348 / residual.cc:235-278.
//
// The reference evaluates about 170 3x3 products for each (cell, q, i, j).
// Because every FESystem(FE_Q(k),18) shape function is one-hot in the
// component, each term is a bilinear form evaluated at unit tensors, so the
// cell matrix factorises exactly as
//   J[(a,c),(b,d)] = sum_q JxW_q N_a N_b H_cd(q)
//                  + K1 delta_cd S_ab + (K2+K3) [same part and spin row] S^{k(c) l(d)}_ab
// with H(q) the 18x18 Hessian of the bulk density at the iterate
// (alpha + 2 sum beta_n m_n) and S the stiffness integrals; the residual
// factorises the same way.  H is evaluated in closed form from six 3x3 Gram
// matrices per quadrature point.
//
// One workgroup (3 wavefronts) owns one cell.  Thread t < 171 owns one
// unique (c <= d) component pair, keeps H_cd(q) for all q in registers and
// contracts it against the (a <= b) node-pair weights, which the workgroup
// stages in LDS 32 pairs at a time (the whole table, 82 KB for Q2, does not
// fit the scalar cache).  It writes entries [c][d] and [d][c] of block (a,b)
// and of its transpose (b,a) straight to the BSR values, one node pair after
// another, with the old values of the next two pairs already requested; the
// block codes and node masks of the cell sit in LDS as well.
// Cells of one colour share no node, so the read-modify-write scatter needs
// no atomics and the result is bit-deterministic; the first colour to touch
// a block stores instead of accumulating, so J needs no zero fill.
// Dirichlet rows/columns are condensed in the scatter (constrained entries
// 0, constrained diagonal += |local diagonal|, the documented behaviour of
// deal.II's AffineConstraints::distribute_local_to_global).
#include "glhip_dev.hpp"

#include <type_traits>

namespace glh {
namespace {

constexpr int NT = 192;   // threads per workgroup (3 wavefronts)
constexpr int GRAM = 57;  // per-q Gram data (see gram_entry)
constexpr int NUPAIR = 171;

// Gram layout per quadrature point:
//  [0,9) Auu = u^T u (orbital k,l)   [9,18) Avv = v^T v   [18,27) Auv = u^T v
//  [27,36) uut = u u^T (spin mu,nu)  [36,45) vvt          [45,54) uvt = u v^T
//  54 |u|^2  55 |v|^2  56 u:v
__device__ inline double gram_entry(const double *x, int g)
{
  const double *u = x, *v = x + 9;
  if (g < 54) {
    const int blk = g / 9, i = (g % 9) / 3, j = g % 3;
    const double *A = (blk == 0 || blk == 2 || blk == 3 || blk == 5) ? u : v;
    const double *B = (blk == 0 || blk == 3) ? u : v;
    double s = 0.0;
    if (blk < 3) {  // A^T B: sum over spin rows
      for (int m = 0; m < 3; ++m) s += A[3 * m + i] * B[3 * m + j];
    } else {        // A B^T: sum over orbital columns
      for (int k = 0; k < 3; ++k) s += A[3 * i + k] * B[3 * j + k];
    }
    return s;
  }
  double s = 0.0;
  if (g == 54) for (int a = 0; a < 9; ++a) s += u[a] * u[a];
  else if (g == 55) for (int a = 0; a < 9; ++a) s += v[a] * v[a];
  else for (int a = 0; a < 9; ++a) s += u[a] * v[a];
  return s;
}

// H_cd (c <= d) of the bulk density: alpha delta + 2 sum_n beta_n m_n(E_c, E_d).
__device__ inline double hess_entry(const double *__restrict__ G, const double *__restrict__ x, int c, int d,
                                    const Material &M)
{
  const double *u = x, *v = x + 9;
  const int pc = c / 9, pd = d / 9, ac = c % 9, ad = d % 9;
  const int mu = ac / 3, k = ac % 3, nu = ad / 3, l = ad % 3;
  const double dmn = (mu == nu) ? 1.0 : 0.0, dkl = (k == l) ? 1.0 : 0.0, dab = dmn * dkl;
  const double U2 = G[54], V2 = G[55], UV = G[56], D = U2 - V2, N = U2 + V2;
  const double b345 = M.b3 + M.b4 + M.b5;
  if (pc == pd) {
    // uu block, or vv block with the roles of u and v exchanged
    const double *p = pc ? v : u, *o = pc ? u : v;
    const double *App = pc ? G + 9 : G, *Aoo = pc ? G : G + 9;
    const double *Spp = pc ? G + 36 : G + 27, *Soo = pc ? G + 27 : G + 36;
    const double sD = pc ? -D : D;
    return dab * (M.al + 2.0 * M.b1 * sD + 2.0 * M.b2 * N) + 4.0 * (M.b1 + M.b2) * p[ac] * p[ad] +
           4.0 * M.b1 * o[ac] * o[ad] +
           2.0 * dmn * (b345 * App[3 * l + k] + (M.b3 + M.b4 - M.b5) * Aoo[3 * l + k]) +
           2.0 * b345 * p[3 * mu + l] * p[3 * nu + k] + 2.0 * (M.b3 - M.b4 + M.b5) * o[3 * mu + l] * o[3 * nu + k] +
           2.0 * dkl * (b345 * Spp[3 * mu + nu] + (-M.b3 + M.b4 + M.b5) * Soo[3 * mu + nu]);
  }
  // uv block: row (mu,k) in u, column (nu,l) in v
  const double Auv_lk = G[18 + 3 * l + k], Avu_lk = G[18 + 3 * k + l];
  const double uvt_mn = G[45 + 3 * mu + nu], uvt_nm = G[45 + 3 * nu + mu];
  const double Q = uvt_mn + uvt_nm, W = uvt_nm - uvt_mn;
  const double ua = u[ac], va = v[ac], ub = u[ad], vb = v[ad];
  const double uml = u[3 * mu + l], unk = u[3 * nu + k], vml = v[3 * mu + l], vnk = v[3 * nu + k];
  return dab * 4.0 * M.b1 * UV + 4.0 * M.b1 * (ub * va - vb * ua) + 4.0 * M.b2 * ua * vb +
         2.0 * M.b3 * (-dmn * Avu_lk - vml * unk + dmn * Auv_lk + uml * vnk + Q * dkl) +
         2.0 * M.b4 * (dmn * Avu_lk + vml * unk - dmn * Auv_lk + uml * vnk - W * dkl) +
         2.0 * M.b5 * (dmn * Avu_lk + vml * unk + dmn * Auv_lk - uml * vnk + W * dkl);
}

// rho_c = alpha x_c + 2 sum_n beta_n r_n(E_c), the bulk residual density, for
// the components (mu, k) of u (ru) and of v (rv) at once, from the 18 values
// x of one point.  The Gram entries it needs (|u|^2, |v|^2, u:v and row mu of
// u u^T, v v^T, u v^T, v u^T) are summed in the order of gram_entry; x sits in
// registers, and its entries at run-time indices are read again from memory
// (LDS) so that no register array is indexed at run time.
template <typename TA>
__device__ inline void rho_pair(const TA *x, int mu, int k, const Material &M, double &ru, double &rv)
{
  double u[9], v[9];
#pragma unroll
  for (int a = 0; a < 9; ++a) { u[a] = (double)x[a]; v[a] = (double)x[9 + a]; }
  double U2 = 0.0, V2 = 0.0, UV = 0.0;
#pragma unroll
  for (int a = 0; a < 9; ++a) U2 += u[a] * u[a];
#pragma unroll
  for (int a = 0; a < 9; ++a) V2 += v[a] * v[a];
#pragma unroll
  for (int a = 0; a < 9; ++a) UV += u[a] * v[a];
  const double D = U2 - V2, N = U2 + V2;
  double um[3], vm[3];  // row mu of u and of v
#pragma unroll
  for (int j = 0; j < 3; ++j) { um[j] = (double)x[3 * mu + j]; vm[j] = (double)x[9 + 3 * mu + j]; }
  double Pu = 0, Pv = 0, Qu = 0, Qv = 0, Su = 0, Sv = 0, Wu = 0, Wv = 0;
#pragma unroll
  for (int nu = 0; nu < 3; ++nu) {
    double uu = 0.0, vv = 0.0, uvmn = 0.0, uvnm = 0.0;  // uut, vvt, uvt at (mu, nu); uvt at (nu, mu)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      uu += um[j] * u[3 * nu + j];
      vv += vm[j] * v[3 * nu + j];
      uvmn += um[j] * v[3 * nu + j];
      uvnm += u[3 * nu + j] * vm[j];
    }
    const double P = uu - vv, S = uu + vv, Q = uvmn + uvnm, W = uvnm - uvmn;
    const double un = (double)x[3 * nu + k], vn = (double)x[9 + 3 * nu + k];
    Pu += P * un; Pv += P * vn; Qu += Q * un; Qv += Q * vn;
    Su += S * un; Sv += S * vn; Wu += W * un; Wv += W * vn;
  }
  const double ua = (double)x[3 * mu + k], va = (double)x[9 + 3 * mu + k];
  ru = M.al * ua + 2.0 * (M.b1 * (D * ua + 2.0 * UV * va) + M.b2 * (N * ua) + M.b3 * (Pu + Qv) + M.b4 * (Su - Wv) +
                          M.b5 * (Su + Wv));
  rv = M.al * va + 2.0 * (M.b1 * (-D * va + 2.0 * UV * ua) + M.b2 * (N * va) + M.b3 * (Qu - Pv) + M.b4 * (Sv + Wu) +
                          M.b5 * (Sv - Wu));
}

// H(s) x for the components (mu, k) of u (zu) and of v (zv) without forming
// H: the derivative of rho at the state s in the direction x (18 values
// each, in LDS), term by term the linearisation of rho_pair:
//   d|u|^2 = 2 u:a, d|v|^2 = 2 v:b, d(u:v) = u:b + a:v,
//   d(u u^T) = a u^T + u a^T, d(v v^T) = b v^T + v b^T, d(u v^T) = a v^T + u b^T,
//   d(P u) = dP u + P a, ...  with (a, b) the u and v parts of x.
// In two stages, so that a point's 9 items share the sums through LDS
// instead of each forming all of them.  The row-pair sums are symmetric in
// (mu, nu) but for W and dW, which change sign:
//   uu, vv symmetric; uvmn(nu, mu) = uvnm(mu, nu), so Q is symmetric and W
//   antisymmetric; likewise duu, dvv, dQ and dW.
// The six pairs mu <= nu, numbered (0,0) (0,1) (0,2) (1,1) (1,2) (2,2), hold
// 8 values each, P S Q W dP dS dQ dW; the pairs (mu, mu) also give the 9-term
// sums: D = sum_mu P, N = sum_mu S, u:v = sum_mu Q / 2, dD = sum_mu dP,
// dN = sum_mu dS, d(u:v) = sum_mu dQ / 2.
// Stage 1: the 8 sums of pair p at the state s in the direction x, to r.
template <typename TA>
__device__ inline void hx_row_sums(const TA *s, const TA *x, int p, TA *r)
{
  const int mu = p < 3 ? 0 : (p < 5 ? 1 : 2), nu = p < 3 ? p : (p < 5 ? p - 2 : 2);
  TA uu = 0, vv = 0, uvmn = 0, uvnm = 0, duu = 0, dvv = 0, duvmn = 0, duvnm = 0;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const TA um = s[3 * mu + j], vm = s[9 + 3 * mu + j], am = x[3 * mu + j], bm = x[9 + 3 * mu + j];
    const TA un = s[3 * nu + j], vn = s[9 + 3 * nu + j], an = x[3 * nu + j], bn = x[9 + 3 * nu + j];
    uu += um * un;
    vv += vm * vn;
    uvmn += um * vn;
    uvnm += un * vm;
    duu += am * un + um * an;
    dvv += bm * vn + vm * bn;
    duvmn += am * vn + um * bn;
    duvnm += an * vm + un * bm;
  }
  r[0] = uu - vv;
  r[1] = uu + vv;
  r[2] = uvmn + uvnm;
  r[3] = uvnm - uvmn;
  r[4] = duu - dvv;
  r[5] = duu + dvv;
  r[6] = duvmn + duvnm;
  r[7] = duvnm - duvmn;
}

// Stage 2: H(s) x for the components (mu, k) of u (zu) and of v (zv) from the
// point's 48 pair sums R (stage 1) and its s and x.
template <typename TA>
__device__ inline void hx_from_sums(const TA *R, const TA *s, const TA *x, int mu, int k, const Material &M, TA &zu,
                                    TA &zv)
{
  // the 9-term sums from the pairs (0,0), (1,1), (2,2)
  TA D = 0, N = 0, Q2 = 0, dD = 0, dN = 0, dQ2 = 0;
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const TA *r = R + 8 * (m == 0 ? 0 : (m == 1 ? 3 : 5));
    D += r[0];
    N += r[1];
    Q2 += r[2];
    dD += r[4];
    dN += r[5];
    dQ2 += r[6];
  }
  const TA UV = (TA)0.5 * Q2, dUV = (TA)0.5 * dQ2;
  TA dPu = 0, dPv = 0, dQu = 0, dQv = 0, dSu = 0, dSv = 0, dWu = 0, dWv = 0;
#pragma unroll
  for (int nu = 0; nu < 3; ++nu) {
    const int lo = mu < nu ? mu : nu, hi = mu < nu ? nu : mu;
    const TA *r = R + 8 * (lo == 0 ? hi : (lo == 1 ? hi + 2 : 5));
    const TA sg = mu > nu ? (TA)-1 : (TA)1;  // W and dW of the pair (mu, nu)
    const TA P = r[0], S = r[1], Q = r[2], W = sg * r[3], dP = r[4], dS = r[5], dQ = r[6], dW = sg * r[7];
    const TA un = s[3 * nu + k], vn = s[9 + 3 * nu + k], an = x[3 * nu + k], bn = x[9 + 3 * nu + k];
    dPu += dP * un + P * an; dPv += dP * vn + P * bn;
    dQu += dQ * un + Q * an; dQv += dQ * vn + Q * bn;
    dSu += dS * un + S * an; dSv += dS * vn + S * bn;
    dWu += dW * un + W * an; dWv += dW * vn + W * bn;
  }
  const TA u0 = s[3 * mu + k], v0 = s[9 + 3 * mu + k], a0 = x[3 * mu + k], b0 = x[9 + 3 * mu + k];
  const TA al = (TA)M.al, b1 = (TA)M.b1, b2 = (TA)M.b2, b3 = (TA)M.b3, b4 = (TA)M.b4, b5 = (TA)M.b5;
  zu = al * a0 + (TA)2 * (b1 * (dD * u0 + D * a0 + (TA)2 * (dUV * v0 + UV * b0)) + b2 * (dN * u0 + N * a0) +
                          b3 * (dPu + dQv) + b4 * (dSu - dWv) + b5 * (dSu + dWv));
  zv = al * b0 + (TA)2 * (b1 * ((TA)2 * (dUV * u0 + UV * a0) - (dD * v0 + D * b0)) + b2 * (dN * v0 + N * b0) +
                          b3 * (dQu - dPv) + b4 * (dSv + dWu) + b5 * (dSv - dWu));
}

__device__ inline bool face_zeroed(int bid, int c) { return (c % 9) % 3 == bid - 2; }

__device__ inline double face_area(int lf, const double *h)
{
  return lf < 2 ? h[1] * h[2] : (lf < 4 ? h[0] * h[2] : h[0] * h[1]);
}

template <typename T> struct Vec2;
template <> struct Vec2<double> { using type = double2; };
template <> struct Vec2<float> { using type = float2; };

template <int DEG, typename TJ, bool HANG = false>
__global__ __launch_bounds__(NT) void k_assemble(RefTables T, MeshDev Mh, Material M0, const double *__restrict__ U,
                                                 const int *__restrict__ cells, TJ *__restrict__ J)
{
  constexpr int N1 = DEG + 1, NS = N1 * N1 * N1, NQ = NS, NPAIR = NS * (NS + 1) / 2;
  __shared__ double sU[NS * 18];
  __shared__ double sX[NQ * 18];
  __shared__ double sGram[NQ * GRAM];
  // per-cell scatter data staged once: block codes and node constraint masks
  __shared__ int sBlk[NS * NS];
  // node-pair weights of the current chunk of pairs (the whole table does not
  // fit the scalar cache)
  constexpr int PCH = 32;
  __shared__ double sWp[PCH * NQ];
  __shared__ double sSp[PCH * 9];
  __shared__ uint32_t sMask[NS];

  const int tid = threadIdx.x;
  const int e = cells[blockIdx.x];
  if (Mh.inst_active && !Mh.inst_active[Mh.cell_inst[e]]) return;  // finished instance of a batch
  const Material M = Mh.mats ? Mh.mats[Mh.cell_mat[e]] : M0;  // per-instance (T, p) in a batch
  const int *cn = Mh.cell_nodes + (size_t)e * NS;
  double h[3], ih[3];
  for (int k = 0; k < 3; ++k) { h[k] = Mh.cell_h[3 * e + k]; ih[k] = 1.0 / h[k]; }
  const double vol = h[0] * h[1] * h[2];
  const uint32_t cface = Mh.cell_face[e];

  // gather the cell's 18 components at its NS nodes
  for (int i = tid; i < NS * 18; i += NT) sU[i] = U[(size_t)cn[i / 18] * 18 + (i % 18)];
  {
    const int *cb = Mh.cell_blk + (size_t)e * NS * NS;
    for (int i = tid; i < NS * NS; i += NT) sBlk[i] = cb[i];
    for (int i = tid; i < NS; i += NT) sMask[i] = Mh.node_mask[cn[i]];
  }
  __syncthreads();
  // the iterate at the quadrature points (s_vector2matrix.cc:116-217)
  for (int i = tid; i < NQ * 18; i += NT) {
    const int q = i / 18, c = i % 18;
    double xv = 0;
    for (int a = 0; a < NS; ++a) xv += T.N[q * NS + a] * sU[a * 18 + c];
    sX[i] = xv;
  }
  __syncthreads();
  for (int i = tid; i < NQ * GRAM; i += NT) {
    const int q = i / GRAM, g = i % GRAM;
    sGram[i] = gram_entry(sX + q * 18, g);
  }
  __syncthreads();

  {
    // Thread t < 171 owns the unique component pair (c <= d).  For every node
    // pair (a <= b) it forms block(a,b)[c][d], block(a,b)[d][c] and writes them,
    // with the transposed block(b,a), straight to the BSR values (no staging
    // of the output, so each wavefront keeps many stores in flight; L2 merges
    // the lanes' 8-byte stores into lines).  All 192 threads load each chunk
    // of node-pair weights into LDS between two barriers.
    const bool owner = tid < NUPAIR;
    int c = 0, d = 0;
    {
      int t = owner ? tid : 0;
      while (t >= 18 - c) { t -= 18 - c; ++c; }
      d = c + t;
    }
    double Hq[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) Hq[q] = hess_entry(sGram + q * GRAM, sX + q * 18, c, d, M);
    const bool grad_cd = (c / 9 == d / 9) && ((c % 9) / 3 == (d % 9) / 3);
    const int kc = c % 3, kd = d % 3;
    const double lap0 = vol * ih[0] * ih[0], lap1 = vol * ih[1] * ih[1], lap2 = vol * ih[2] * ih[2];
    const double k23_cd = M.K23 * vol * ih[kc] * ih[kd];
    const uint32_t *mask = Mh.node_mask;
    const int *cblk = sBlk;
    const int o_cd = c * 18 + d, o_dc = d * 18 + c;
    // Robin diagonal weight of this component on each local face (0 if off)
    double rob[6];
#pragma unroll
    for (int lf = 0; lf < 6; ++lf) {
      const int bid = (int)((cface >> (3 * lf)) & 7u) - 1;
      rob[lf] = (c == d && M.robin != 0.0 && bid >= 2 && bid <= 4 && !face_zeroed(bid, c))
                    ? M.robin * face_area(lf, h) : 0.0;
    }
    // Old values of blocks shared with earlier colours are loaded one node pair
    // ahead, so their latency overlaps the contraction of the current pair.
    auto blk_ptr = [&](int code) -> TJ * { return J + (size_t)(code < 0 ? -code - 1 : code) * 324; };
    auto prefetch = [&](int p, double *o) {
      const int a = T.pa[p], b = T.pb[p];
      const int cab = cblk[a * NS + b];
      o[0] = o[1] = o[2] = o[3] = 0.0;
      if (cab >= 0) {
        const TJ *B = J + (size_t)cab * 324;
        o[0] = (double)B[o_cd];
        if (c != d) o[1] = (double)B[o_dc];
      }
      if (a != b) {
        const int cba = cblk[b * NS + a];
        if (cba >= 0) {
          const TJ *B = J + (size_t)cba * 324;
          o[2] = (double)B[o_dc];
          if (c != d) o[3] = (double)B[o_cd];
        }
      }
    };
    // entries [c][d] (e1) and [d][c] (e2) of the local block of node pair p
    auto pair_values = [&](int p, const double *wp, const double *sp, double &e1, double &e2) {
      double sym = 0.0;
#pragma unroll
      for (int q = 0; q < NQ; ++q) sym += wp[q] * Hq[q];
      sym *= vol;
      const int a = T.pa[p], b = T.pb[p];
      if (c == d) {
        sym += M.K1 * (sp[0] * lap0 + sp[4] * lap1 + sp[8] * lap2);
#pragma unroll
        for (int lf = 0; lf < 6; ++lf)
          if (rob[lf] != 0.0) sym += rob[lf] * T.Mf[(lf * NS + a) * NS + b];
      }
      e1 = sym;
      e2 = sym;
      if (grad_cd) {
        e1 += k23_cd * sp[kc * 3 + kd];
        e2 += k23_cd * sp[kd * 3 + kc];
      }
    };
    auto process = [&](int p, const double *ol) {
      double e1, e2;
      pair_values(p, sWp + (p % PCH) * NQ, sSp + (p % PCH) * 9, e1, e2);
      const int a = T.pa[p], b = T.pb[p];
      const uint32_t ma = sMask[a], mb = sMask[b];
      {  // block (a,b): rows of node a, columns of node b
        TJ *B = blk_ptr(cblk[a * NS + b]);
        double v1 = e1, v2 = e2;
        const bool r1 = (ma >> c) & 1u, c1 = (mb >> d) & 1u, r2 = (ma >> d) & 1u, c2 = (mb >> c) & 1u;
        if (r1 || c1) v1 = (a == b && r1 && c == d) ? fabs(v1) : 0.0;
        if (r2 || c2) v2 = (a == b && r2 && c == d) ? fabs(v2) : 0.0;
        B[o_cd] = (TJ)(ol[0] + v1);
        if (c != d) B[o_dc] = (TJ)(ol[1] + v2);
      }
      if (a != b) {  // block (b,a) = block(a,b)^T: [d][c] = e1, [c][d] = e2
        TJ *B = blk_ptr(cblk[b * NS + a]);
        double v1 = e1, v2 = e2;
        const bool r1 = (mb >> d) & 1u, c1 = (ma >> c) & 1u, r2 = (mb >> c) & 1u, c2 = (ma >> d) & 1u;
        if (r1 || c1) v1 = 0.0;
        if (r2 || c2) v2 = 0.0;
        B[o_dc] = (TJ)(ol[2] + v1);
        if (c != d) B[o_cd] = (TJ)(ol[3] + v2);
      }
    };
    // PD node pairs in flight: the old values of pair p + PD are requested
    // before pair p is contracted and stored (PD = 2 measured fastest with
    // the block codes and masks in LDS; the registers stay at 3 waves/SIMD)
    constexpr int PD = 2;
    double ob[PD][4];
    if (owner) {
#pragma unroll
      for (int k = 0; k < PD; ++k) prefetch(k < NPAIR ? k : NPAIR - 1, ob[k]);
    }
    for (int c0 = 0; c0 < NPAIR; c0 += PCH) {
      const int np = NPAIR - c0 < PCH ? NPAIR - c0 : PCH;
      __syncthreads();  // the previous chunk has been used
      for (int i = tid; i < np * NQ; i += NT) sWp[i] = T.Wp[(size_t)c0 * NQ + i];
      for (int i = tid; i < np * 9; i += NT) sSp[i] = T.Sp[(size_t)c0 * 9 + i];
      __syncthreads();
      if (!owner) continue;
      for (int p0 = c0; p0 < c0 + np; p0 += PD) {
#pragma unroll
        for (int k = 0; k < PD; ++k) {
          const int p = p0 + k;
          if (p < c0 + np) {
            double ol[4] = {ob[k][0], ob[k][1], ob[k][2], ob[k][3]};
            if (p + PD < NPAIR) prefetch(p + PD, ob[k]);
            process(p, ol);
          }
        }
      }
    }
    if (!owner) return;
    // hanging nodes (adaptive meshes only): the condensed row keeps the sum
    // over cells of |local diagonal|, gathered outside the hot loop
    if (HANG && c == d)
      for (int p = 0; p < NPAIR; ++p) {
        const int a = T.pa[p];
        if (a != T.pb[p] || !(mask[cn[a]] & HANG_BIT)) continue;
        double e1, e2;
        pair_values(p, T.Wp + p * NQ, T.Sp + p * 9, e1, e2);
        Mh.hdiag[(size_t)cn[a] * 18 + c] += fabs(e1);
      }
  }
}

// ------------------------------------------------ matrix-free Jacobian
// J x without the assembled matrix (on a mesh with constraint lines applied
// to C x and condensed, gl_constraints.hip).  The Jacobian of the cell loop is
//   J_(a,c),(b,d) = vol sum_q w_q N_a N_b H_cd(q)             (bulk, H symmetric)
//                 + K1 delta_cd vol sum_q w_q grad N_a . grad N_b
//                 + K23 [c, d same part and spin] vol sum_q w_q d_kc N_a d_kd N_b
//                 + Robin face mass on the diagonal components,
// with kc, kd the orbital indices of c, d.  Per cell this is applied as
//   X_q = sum_b N_b(q) x_b, G_kq = sum_b d_k N_b(q) x_b / h_k,
//   y_a += sum_q w_q vol (N_a(q) H(q) X_q + sum_k d_k N_a(q) / h_k gamma_k(q)),
//   gamma_{c,k} = K1 G_kq[c] + K23 [k == kc] sum_kd G_kd,q[9 part + 3 mu + kd],
// sum-factorised, instead of streaming the cell's 729 blocks.  k_mf_setup
// stores the iterate at the quadrature points per cell (the only state of
// the operator besides the diagonal blocks; H(q) X(q) is formed from it in
// the product) and assembles the diagonal blocks for the point-block Jacobi
// preconditioner with the arithmetic of k_assemble; Dirichlet columns are
// applied as a zero x there, Dirichlet rows keep only their diagonal
// (k_mf_gather).

// CON: a mesh with constraint lines (the hanging nodes' |local diagonal| and
// the raw local blocks of X as well)
template <int DEG, typename TH, bool CON>
__global__ __launch_bounds__(NT) void k_mf_setup(RefTables T, MeshDev Mh, Material M0, const double *__restrict__ U,
                                                 const int *__restrict__ cells, TH *__restrict__ Uq,
                                                 double *__restrict__ Jd, MfCondDev X)
{
  constexpr int N1 = DEG + 1, NS = N1 * N1 * N1, NQ = NS;
  __shared__ double sU[NS * 18];
  __shared__ double sX[NQ * 18];
  __shared__ double sGram[NQ * GRAM];
  // per node of the cell: global node, constraint mask, and whether this
  // cell writes the node's diagonal block first
  __shared__ int sNode[NS];
  __shared__ uint32_t sMask[NS];
  __shared__ int sFirst[NS];
  const int tid = threadIdx.x;
  const int e = cells[blockIdx.x];
  // a finished instance of a batch keeps the iterate and blocks of its last
  // step (all of its cells are skipped together)
  if (Mh.inst_active && !Mh.inst_active[Mh.cell_inst[e]]) return;
  const Material M = Mh.mats ? Mh.mats[Mh.cell_mat[e]] : M0;
  const int *cn = Mh.cell_nodes + (size_t)e * NS;
  double h[3], ih[3];
  for (int k = 0; k < 3; ++k) { h[k] = Mh.cell_h[3 * e + k]; ih[k] = 1.0 / h[k]; }
  const double vol = h[0] * h[1] * h[2];
  const uint32_t cface = Mh.cell_face[e];
  const int *cblk = Mh.cell_blk + (size_t)e * NS * NS;
  for (int i = tid; i < NS * 18; i += NT) sU[i] = U[(size_t)cn[i / 18] * 18 + (i % 18)];
  if (tid < NS) {
    const int node = cn[tid];
    sNode[tid] = node;
    sMask[tid] = Mh.node_mask[node];
    sFirst[tid] = cblk[tid * NS + tid] < 0;
  }
  __syncthreads();
  for (int i = tid; i < NQ * 18; i += NT) {
    const int q = i / 18, c = i % 18;
    double xv = 0;
    for (int a = 0; a < NS; ++a) xv += T.N[q * NS + a] * sU[a * 18 + c];
    sX[i] = xv;
    Uq[(size_t)e * NQ * 18 + i] = (TH)xv;  // the iterate at the points, the product's state
  }
  __syncthreads();
  for (int i = tid; i < NQ * GRAM; i += NT) {
    const int q = i / GRAM, g = i % GRAM;
    sGram[i] = gram_entry(sX + q * 18, g);
  }
  __syncthreads();
  if (tid >= NUPAIR) return;
  int c = 0, d = 0;
  {
    int t = tid;
    while (t >= 18 - c) { t -= 18 - c; ++c; }
    d = c + t;
  }
  double Hv[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) Hv[q] = hess_entry(sGram + q * GRAM, sX + q * 18, c, d, M);
  // diagonal blocks (a,a), as k_assemble forms and masks them
  const bool grad_cd = (c / 9 == d / 9) && ((c % 9) / 3 == (d % 9) / 3);
  const int kc = c % 3, kd = d % 3;
  const double lap0 = vol * ih[0] * ih[0], lap1 = vol * ih[1] * ih[1], lap2 = vol * ih[2] * ih[2];
  const double k23_cd = M.K23 * vol * ih[kc] * ih[kd];
  const int o_cd = c * 18 + d, o_dc = d * 18 + c;
  // entries [c][d] (v1) and [d][c] (v2) of the local block of node pair p =
  // (a <= b), before the Dirichlet rows and columns are applied
  auto pair_block = [&](int p, int a, int b, double &v1, double &v2) {
    const double *wp = T.Wp + (size_t)p * NQ, *sp = T.Sp + (size_t)p * 9;
    double sym = 0.0;
#pragma unroll
    for (int q = 0; q < NQ; ++q) sym += wp[q] * Hv[q];
    sym *= vol;
    if (c == d) {
      sym += M.K1 * (sp[0] * lap0 + sp[4] * lap1 + sp[8] * lap2);
      if (M.robin != 0.0)
        for (int lf = 0; lf < 6; ++lf) {
          const int bid = (int)((cface >> (3 * lf)) & 7u) - 1;
          if (bid >= 2 && bid <= 4 && !face_zeroed(bid, c))
            sym += (M.robin * face_area(lf, h)) * T.Mf[(lf * NS + a) * NS + b];
        }
    }
    v1 = sym;
    v2 = sym;
    if (grad_cd) {
      v1 += k23_cd * sp[kc * 3 + kd];
      v2 += k23_cd * sp[kd * 3 + kc];
    }
  };
  // The nodes' blocks are updated a chunk of nodes at a time: the chunk's
  // contributions, then all of its old values requested at once, then the
  // stores, so that one load latency covers the chunk (the nodes of a cell
  // are distinct, so the chunk's blocks do not overlap).
  constexpr int AC = DEG == 1 ? 4 : 3;
  static_assert(NS % AC == 0, "whole chunks of nodes");
  for (int a0 = 0; a0 < NS; a0 += AC) {
    double v1[AC], v2[AC], o1[AC], o2[AC];
#pragma unroll
    for (int k = 0; k < AC; ++k) {
      const int a = a0 + k;
      pair_block(a * NS - a * (a - 1) / 2, a, a, v1[k], v2[k]);  // pair (a, a)
      const uint32_t ma = sMask[a];
      // a hanging node's condensed row keeps the sum over cells of |local diagonal|
      if (CON && c == d && (ma & HANG_BIT)) Mh.hdiag[(size_t)sNode[a] * 18 + c] += fabs(v1[k]);
      const bool r1 = (ma >> c) & 1u, c1 = (ma >> d) & 1u, r2 = (ma >> d) & 1u, c2 = (ma >> c) & 1u;
      if (r1 || c1) v1[k] = (r1 && c == d) ? fabs(v1[k]) : 0.0;
      if (r2 || c2) v2[k] = (r2 && c == d) ? fabs(v2[k]) : 0.0;
    }
#pragma unroll
    for (int k = 0; k < AC; ++k) {
      const double *B = Jd + (size_t)sNode[a0 + k] * 324;
      const bool first = sFirst[a0 + k];
      o1[k] = first ? 0.0 : B[o_cd];
      o2[k] = (first || c == d) ? 0.0 : B[o_dc];
    }
#pragma unroll
    for (int k = 0; k < AC; ++k) {
      double *B = Jd + (size_t)sNode[a0 + k] * 324;
      B[o_cd] = o1[k] + v1[k];
      if (c != d) B[o_dc] = o2[k] + v2[k];
    }
  }
  // constrained meshes: the raw local blocks that constraint lines fold into
  // the condensed diagonal blocks of masters (few cells have any)
  if (CON)
    for (int k = X.xptr[e]; k < X.xptr[e + 1]; ++k) {
      const int p = X.xpair[k];
      double v1, v2;
      pair_block(p, T.pa[p], T.pb[p], v1, v2);
      double *B = X.LB + (size_t)k * 324;
      B[o_cd] = v1;
      if (c != d) B[o_dc] = v2;
    }
}

// The product in sum-factorised form: N_b(q) = L_bx(x_qx) L_by(x_qy) L_bz(x_qz),
// so x and its gradients at the quadrature points take three 1D passes (x, y,
// z) and the test against N_a and its gradients three more (z, y, x),
// 2 x 18 x 27 x 27 multiply-adds per Q2 cell instead of 8 x 18 x 27 x 27.
// One thread per 1D line of points and component (a pencil) in each pass: it
// reads its N1 inputs once and applies the 1D bases from registers, so the
// passes move about 2.5x fewer LDS words than one thread per (point,
// component) with the bases in LDS.  At the points, H(q) X(q) is not read
// from a stored Hessian (171 values per point) but formed as the derivative
// of rho at the stored iterate U(q) (18 values per point) in the direction
// X(q) (hx_row_sums, hx_from_sums): the operator's state shrinks from
// 1.21 GB to 0.13 GB on the 32^3 Q2 cube, and the kernel is bound by its
// arithmetic and LDS passes instead of the Hessian stream.  At the points a
// thread owns the components c and c + 9 of one point (the same orbital
// entry in u and in v), after six threads per point have formed the point's
// row-pair sums.
// TH: the stored iterate (FP64, or FP32 for the FP32 inner solve); TX: x.
// TA, the arithmetic, LDS and cell-slot type, is FP32 when both are FP32 (the
// FP32 inner solve, with FP32 accumulation) and FP64 otherwise.  cells[]
// lists the cells in the launch order, -1 for padding (the XCD-aware order of
// mf_cell_order).
// RES: the residual R(x) of the cell loop (residual.cc:164-232, the Robin
// faces :235-278) through the same passes: x is the iterate (no column
// masking), and rho(X(q)), the bulk residual density, takes the place of
// H(q) X(q); Uq is not read, and flag is the batch's instance_active.
template <int DEG, typename TH>
constexpr int mf_threads() { return DEG == 1 ? 128 : 256; }
template <typename TH, typename TX>
using mf_arith = typename std::conditional<std::is_same<TH, float>::value && std::is_same<TX, float>::value, float,
                                           double>::type;

// waves per SIMD the register allocation aims at (FP64: also the LDS limit of
// 4 workgroups per CU)
template <int DEG, typename TH, typename TX>
constexpr int mf_waves() { return std::is_same<mf_arith<TH, TX>, float>::value ? (DEG == 1 ? 5 : 6) : 4; }

template <int DEG, typename TH, typename TX, bool RES = false>
__global__ __launch_bounds__((mf_threads<DEG, TH>()), (mf_waves<DEG, TH, TX>())) void k_mf_apply(RefTables T, MeshDev Mh, Material M0,
                                                                      const TH *__restrict__ Uq,
                                                                      const TX *__restrict__ x,
                                                                      const int *__restrict__ cells,
                                                                      const int *__restrict__ flag,
                                                                      mf_arith<TH, TX> *__restrict__ ycell)
{
  static_assert(!RES || (std::is_same<TH, double>::value && std::is_same<TX, double>::value), "FP64 residual");
  using TA = mf_arith<TH, TX>;
  constexpr int N1 = DEG + 1, NS = N1 * N1 * N1, NQ = NS, V = NS * 18, NTA = mf_threads<DEG, TH>();
  constexpr int NP = N1 * N1 * 18;  // pencils per pass
  constexpr int NH = NQ * 9;        // (point, component pair) items of H X
  static_assert(NH <= NTA, "one H X item per thread");
  // Regions of one value per (point or node, component), in the order
  //   x -> A0 A1 -> B0 B1 B2 -> X G0 G1 G2 -> Z K0 K1 K2 -> C0 C1 C2 -> D0 D1,
  // and the iterate U(q) of the product.
  // A pencil reads exactly the points it writes, so the FP32 kernel lets each
  // pass overwrite its inputs in place in one set of four regions, and keeps
  // the point stage's pair sums and Z in four more (17.7 KB for Q2 with U(q):
  // register-limited at 6 waves per SIMD).  The FP64 kernel alternates
  // between two sets and keeps x (39 KB, 4 workgroups per CU), which saves a
  // barrier and the Robin faces' second gather of x.
  constexpr bool IP = std::is_same<TA, float>::value;
  constexpr int NXR = (V + NTA - 1) / NTA;  // x values per thread
  // room for the unconditional x and U(q) stores, U(q) after x
  constexpr int SU_OFF = (IP ? 3 * V : 8 * V) + NXR * NTA;
  constexpr int RS_OFF = SU_OFF + (RES ? 0 : NXR * NTA);
  __shared__ TA S[RS_OFF + (IP && !RES ? 4 * V : 0)];
  TA *const P0 = S, *const P1 = IP ? S : S + 4 * V, *const SX = IP ? S : S + 8 * V, *const SU = S + SU_OFF;
  const int tid = threadIdx.x;
  const int e = cells[blockIdx.x];
  if (e < 0) return;                           // padding of the XCD-aware order
  if (flag && !flag[Mh.cell_inst[e]]) return;  // instance not iterating: its rows are not read
  const Material M = Mh.mats ? Mh.mats[Mh.cell_mat[e]] : M0;
  const int *cn = Mh.cell_nodes + (size_t)e * NS;
  double h[3];
  TA ih[3];
  for (int k = 0; k < 3; ++k) { h[k] = Mh.cell_h[3 * e + k]; ih[k] = (TA)(1.0 / h[k]); }
  const double vol = h[0] * h[1] * h[2];
  const uint32_t cface = Mh.cell_face[e];
  TA L[N1][N1], dL[N1][N1];  // L[q][b] = L_b(x_q): uniform, scalar loads into scalar registers
  const TA *tL, *tdL;
  if constexpr (std::is_same<TA, float>::value) { tL = T.L1f; tdL = T.dL1f; }
  else { tL = T.L1; tdL = T.dL1; }
#pragma unroll
  for (int q = 0; q < N1; ++q)
#pragma unroll
    for (int b = 0; b < N1; ++b) {
      L[q][b] = tL[q * N1 + b];
      dL[q][b] = tdL[q * N1 + b];
    }
  // A wave's vector loads may return out of order, so a wait for one is a
  // wait for all outstanding.  x at the cell's nodes is therefore gathered
  // and stored to LDS first, and U(q) is requested after that wait: it stays
  // in flight through the interpolation passes, which wait only on LDS.
  // Both without branches (indices clamped into range, unused values stored
  // to padding).
  // x, Dirichlet components 0 (the masked columns; the residual's iterate as it is):
  auto xval = [&](int i) -> TA {
    const int node = cn[i / 18], c = i % 18;
    const TA xv = (TA)x[(size_t)node * 18 + c];
    if constexpr (RES) return xv;
    return ((Mh.node_mask[node] >> c) & 1u) ? (TA)0 : xv;
  };
#pragma unroll
  for (int r = 0; r < NXR; ++r) {
    const int i = tid + r * NTA;
    const TA v = xval(i < V ? i : V - 1);
    SX[i] = i < V ? v : (TA)0;  // beyond V: the padding, never read
  }
  // keep the U(q) loads below the x stores (and so below their wait)
  __builtin_amdgcn_sched_barrier(0);
  TH us[NXR];  // this thread's share of the cell's U(q), stored to LDS after the z pass
  if constexpr (!RES) {
    const TH *Uc = Uq + (size_t)e * V;
#pragma unroll
    for (int r = 0; r < NXR; ++r) {
      const int i = tid + r * NTA;
      us[r] = Uc[i < V ? i : V - 1];
    }
  }
  const int hq = tid / 9 < NQ ? tid / 9 : NQ - 1, hc = tid % 9;  // H X item: point hq, components hc, hc + 9
  __syncthreads();
  auto at = [](int i0, int i1, int i2, int c) { return ((i2 * N1 + i1) * N1 + i0) * 18 + c; };
  const int c = tid % 18, l = tid / 18, la = l % N1, lb = l / N1;  // pencil: line (la, lb), component c
  const bool pen = tid < NP;
  // x pass: line (by, bz) -> (qx, by, bz): x -> A0 A1
  if (pen) {
    TA v[N1];
#pragma unroll
    for (int b = 0; b < N1; ++b) v[b] = SX[at(b, la, lb, c)];
#pragma unroll
    for (int q = 0; q < N1; ++q) {
      TA a0 = 0, a1 = 0;
#pragma unroll
      for (int b = 0; b < N1; ++b) {
        a0 += L[q][b] * v[b];
        a1 += dL[q][b] * v[b];
      }
      P1[at(q, la, lb, c)] = a0;
      P1[V + at(q, la, lb, c)] = a1;
    }
  }
  __syncthreads();
  // y pass: line (qx, bz) -> (qx, qy, bz): A0 A1 -> B0 B1 B2
  if (pen) {
    TA v0[N1], v1[N1];
#pragma unroll
    for (int b = 0; b < N1; ++b) {
      v0[b] = P1[at(la, b, lb, c)];
      v1[b] = P1[V + at(la, b, lb, c)];
    }
#pragma unroll
    for (int q = 0; q < N1; ++q) {
      TA b0 = 0, b1 = 0, b2 = 0;
#pragma unroll
      for (int b = 0; b < N1; ++b) {
        b0 += L[q][b] * v0[b];
        b1 += L[q][b] * v1[b];
        b2 += dL[q][b] * v0[b];
      }
      P0[at(la, q, lb, c)] = b0;
      P0[V + at(la, q, lb, c)] = b1;
      P0[2 * V + at(la, q, lb, c)] = b2;
    }
  }
  __syncthreads();
  // z pass: line (qx, qy) -> (qx, qy, qz) = q: B0 B1 B2 -> X G0 G1 G2
  if (pen) {
    TA v0[N1], v1[N1], v2[N1];
#pragma unroll
    for (int b = 0; b < N1; ++b) {
      v0[b] = P0[at(la, lb, b, c)];
      v1[b] = P0[V + at(la, lb, b, c)];
      v2[b] = P0[2 * V + at(la, lb, b, c)];
    }
#pragma unroll
    for (int q = 0; q < N1; ++q) {
      TA xv = 0, g0 = 0, g1 = 0, g2 = 0;
#pragma unroll
      for (int b = 0; b < N1; ++b) {
        xv += L[q][b] * v0[b];
        g0 += L[q][b] * v1[b];
        g1 += L[q][b] * v2[b];
        g2 += dL[q][b] * v0[b];
      }
      const int o = at(la, lb, q, c);
      P1[o] = xv;
      P1[V + o] = g0 * ih[0];
      P1[2 * V + o] = g1 * ih[1];
      P1[3 * V + o] = g2 * ih[2];
    }
  }
  if constexpr (!RES) {
#pragma unroll
    for (int r = 0; r < NXR; ++r) SU[tid + r * NTA] = (TA)us[r];  // beyond V: padding, never read
  }
  __syncthreads();
  // w vol H(q) X(q) and w vol gamma_k(q) / h_k for the components hc and hc + 9
  // of point hq: read everything, then write Z and K.
  // SPLIT (the product): H(q) X(q) from the point's shared pair sums
  // (hx_row_sums, hx_from_sums), one more barrier.  Regions: the z pass left
  // X G0 G1 G2 in P1; the pair sums (54 per point, 3V in all) go to RS, then
  // Z to Zs = RS + 3V and K over G (Ks = P1 + V); the test passes write C to
  // P0 and D to P1.  FP64: RS is P0, dead after the z pass; FP32: RS is a
  // region of its own, and P0 = P1.  The residual (not SPLIT) writes Z and K
  // over the regions of P0.
  constexpr bool SPLIT = !RES;
  TA *const RS = IP ? S + RS_OFF : P0;
  TA *const Zs = SPLIT ? RS + 3 * V : P0, *const Ks = SPLIT ? P1 + V : P0 + V;
  TA *const Cd = SPLIT ? P0 : P1, *const Dd = SPLIT ? P1 : P0;
  TA zo[2], ko[2][3];
  if (tid < NH) {
    const TA wv = (TA)(T.w[hq] * vol), K1 = (TA)M.K1, K23 = (TA)M.K23;
    const int mu = hc / 3, kc = hc % 3;
    if constexpr (RES) {  // rho first: fewer registers live through it
      double ru, rv;
      rho_pair(&P1[hq * 18], mu, kc, M, ru, rv);
      zo[0] = wv * ru;
      zo[1] = wv * rv;
    }
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int i = hq * 18 + hc + 9 * r, base = hq * 18 + 9 * r + 3 * mu;
      const TA dv = P1[V + base] + P1[2 * V + base + 1] + P1[3 * V + base + 2];
#pragma unroll
      for (int k = 0; k < 3; ++k) ko[r][k] = wv * ih[k] * (K1 * P1[(1 + k) * V + i] + (k == kc ? K23 * dv : (TA)0));
    }
  }
  if constexpr (SPLIT) {
    // pair p of point tid / 6 (162 items for Q2), then the 9 items per point;
    // 54 values per point (48 used) keep each pair's 8 values 16-byte aligned
    // for 128-bit LDS accesses; of the strides 48-54 it ran fastest (an odd
    // stride, 49, cost 10% of the kernel in bank conflicts)
    if (tid < NQ * 6) {
      const int sq = tid / 6, p = tid % 6;
      hx_row_sums(SU + sq * 18, P1 + sq * 18, p, RS + sq * 54 + p * 8);
    }
    __syncthreads();
    if (tid < NH) {
      TA z[2];
      hx_from_sums(RS + hq * 54, SU + hq * 18, P1 + hq * 18, hc / 3, hc % 3, M, z[0], z[1]);
      const TA wv = (TA)(T.w[hq] * vol);
      zo[0] = wv * z[0];
      zo[1] = wv * z[1];
    }
  }
  if constexpr (IP && !SPLIT) __syncthreads();
  if (tid < NH) {
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const int i = hq * 18 + hc + 9 * r;
      Zs[i] = zo[r];
#pragma unroll
      for (int k = 0; k < 3; ++k) Ks[k * V + i] = ko[r][k];
    }
  }
  __syncthreads();
  // test z pass: line (qx, qy) -> (qx, qy, az): Z K0 K1 K2 -> C0 C1 C2
  if (pen) {
    TA z[N1], k0[N1], k1[N1], k2[N1];
#pragma unroll
    for (int q = 0; q < N1; ++q) {
      const int o = at(la, lb, q, c);
      z[q] = Zs[o];
      k0[q] = Ks[o];
      k1[q] = Ks[V + o];
      k2[q] = Ks[2 * V + o];
    }
#pragma unroll
    for (int a = 0; a < N1; ++a) {
      TA c0 = 0, c1 = 0, c2 = 0;
#pragma unroll
      for (int q = 0; q < N1; ++q) {
        c0 += L[q][a] * z[q] + dL[q][a] * k2[q];
        c1 += L[q][a] * k0[q];
        c2 += L[q][a] * k1[q];
      }
      const int o = at(la, lb, a, c);
      Cd[o] = c0;
      Cd[V + o] = c1;
      Cd[2 * V + o] = c2;
    }
  }
  __syncthreads();
  // test y pass: line (qx, az) -> (qx, ay, az): C0 C1 C2 -> D0 D1
  if (pen) {
    TA v0[N1], v1[N1], v2[N1];
#pragma unroll
    for (int q = 0; q < N1; ++q) {
      const int o = at(la, q, lb, c);
      v0[q] = Cd[o];
      v1[q] = Cd[V + o];
      v2[q] = Cd[2 * V + o];
    }
#pragma unroll
    for (int a = 0; a < N1; ++a) {
      TA d0 = 0, d1 = 0;
#pragma unroll
      for (int q = 0; q < N1; ++q) {
        d0 += L[q][a] * v0[q] + dL[q][a] * v2[q];
        d1 += L[q][a] * v1[q];
      }
      Dd[at(la, a, lb, c)] = d0;
      Dd[V + at(la, a, lb, c)] = d1;
    }
  }
  __syncthreads();
  // test x pass: line (ay, az) -> (ax, ay, az) = a, then the Robin faces
  if (pen) {
    TA v0[N1], v1[N1];
#pragma unroll
    for (int q = 0; q < N1; ++q) {
      v0[q] = Dd[at(q, la, lb, c)];
      v1[q] = Dd[V + at(q, la, lb, c)];
    }
#pragma unroll
    for (int a = 0; a < N1; ++a) {
      TA r = 0;
#pragma unroll
      for (int q = 0; q < N1; ++q) r += L[q][a] * v0[q] + dL[q][a] * v1[q];
      const int na = (lb * N1 + la) * N1 + a, o = at(a, la, lb, c);
      if (M.robin != 0.0)
        for (int lf = 0; lf < 6; ++lf) {
          const int bid = (int)((cface >> (3 * lf)) & 7u) - 1;
          if (bid < 2 || bid > 4 || face_zeroed(bid, c)) continue;
          const double *mf = T.Mf + (lf * NS + na) * NS;
          TA sf = 0;
          for (int b = 0; b < NS; ++b) sf += (TA)mf[b] * (IP ? xval(b * 18 + c) : SX[b * 18 + c]);
          r += (TA)(M.robin * face_area(lf, h)) * sf;
        }
      ycell[(size_t)e * V + o] = r;  // Dirichlet rows are replaced in the gather
    }
  }
}

// The sums over a node's cell slots of components c and c + 1, in the
// node's cell order: four slot indices, then their four loads in flight.
template <typename TY>
__device__ __forceinline__ void slot_pair_sum(const int *__restrict__ ptr, const int *__restrict__ idx,
                                              const TY *__restrict__ ycell, int node, int c, TY &s0, TY &s1)
{
  using Y2 = typename Vec2<TY>::type;
  const int k1 = ptr[node + 1];
  int k = ptr[node];
  for (; k + 4 <= k1; k += 4) {
    int j[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) j[u] = idx[k + u];
    Y2 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = *reinterpret_cast<const Y2 *>(ycell + (size_t)j[u] * 18 + c);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      s0 += v[u].x;
      s1 += v[u].y;
    }
  }
  for (; k < k1; ++k) {
    const Y2 v = *reinterpret_cast<const Y2 *>(ycell + (size_t)idx[k] * 18 + c);
    s0 += v.x;
    s1 += v.y;
  }
}

// y = sum over the cells of each node of their contributions (fixed cell
// order), and for the Dirichlet rows their diagonal entry (the sum over cells
// of |local diagonal|) times x; with b != null the residual y = b - J x.  One
// thread per node and component pair (c, c + 1): two slot values per load.
// TY, the slot and accumulation type, is that of the product (FP32 for the
// FP32 inner solve).
template <typename TX, typename TY>
__global__ void k_mf_gather(int n_nodes, const int *__restrict__ ptr, const int *__restrict__ idx,
                            const TY *__restrict__ ycell, const uint32_t *__restrict__ mask,
                            const double *__restrict__ Jd, const double *__restrict__ sh, const TX *__restrict__ x,
                            const TX *__restrict__ b, const int *__restrict__ row_seg, const int *__restrict__ flag,
                            TX *__restrict__ y)
{
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)n_nodes * 9) return;
  const int node = (int)(i / 9), c = 2 * (int)(i % 9);
  if (flag && !flag[row_seg[node]]) return;  // instance not iterating: its rows are left as they are
  TY s0 = 0, s1 = 0;
  slot_pair_sum(ptr, idx, ycell, node, c, s0, s1);
  const uint32_t mk = mask[node];
  const size_t o = (size_t)node * 18 + c;
  double r0 = ((mk >> c) & 1u) ? Jd[(size_t)node * 324 + c * 19] * (double)x[o] : (double)s0;
  double r1 = ((mk >> (c + 1)) & 1u) ? Jd[(size_t)node * 324 + (c + 1) * 19] * (double)x[o + 1] : (double)s1;
  if (sh) {  // zero on the Dirichlet and hanging rows
    r0 += sh[o] * (double)x[o];
    r1 += sh[o + 1] * (double)x[o + 1];
  }
  y[o] = b ? (TX)((double)b[o] - r0) : (TX)r0;
  y[o + 1] = b ? (TX)((double)b[o + 1] - r1) : (TX)r1;
}

// system_rhs = -R from the residual's cell contributions: each node's sum in
// the fixed cell order, negated; 0 on the Dirichlet components and on the
// rows of a batch instance that is not active (whose cells were skipped).
// Every row is written, so the output needs no fill.
__global__ void k_res_gather(int n_nodes, const int *__restrict__ ptr, const int *__restrict__ idx,
                             const double *__restrict__ ycell, const uint32_t *__restrict__ mask,
                             const int *__restrict__ row_seg, const int *__restrict__ active,
                             double *__restrict__ rhs)
{
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)n_nodes * 9) return;
  const int node = (int)(i / 9), c = 2 * (int)(i % 9);
  const bool on = !active || active[row_seg[node]];
  double s0 = 0.0, s1 = 0.0;
  if (on) slot_pair_sum(ptr, idx, ycell, node, c, s0, s1);
  const uint32_t mk = mask[node];
  const size_t o = (size_t)node * 18 + c;
  rhs[o] = (on && !((mk >> c) & 1u)) ? -s0 : 0.0;
  rhs[o + 1] = (on && !((mk >> (c + 1)) & 1u)) ? -s1 : 0.0;
}

template <int DEG>
__global__ __launch_bounds__(NT) void k_energy(RefTables T, MeshDev Mh, Material M0, const double *__restrict__ U,
                                               int n_cells, double *__restrict__ Fcell)
{
  constexpr int N1 = DEG + 1, NS = N1 * N1 * N1, NQ = NS;
  __shared__ double sU[NS * 18];
  __shared__ double sX[NQ * 18];
  __shared__ double sG[NQ * 54];
  __shared__ double sF[NT];
  const int tid = threadIdx.x, e = blockIdx.x;
  if (e >= n_cells) return;
  const Material M = Mh.mats ? Mh.mats[Mh.cell_mat[e]] : M0;
  const int *cn = Mh.cell_nodes + (size_t)e * NS;
  double h[3], ih[3];
  for (int k = 0; k < 3; ++k) { h[k] = Mh.cell_h[3 * e + k]; ih[k] = 1.0 / h[k]; }
  const double vol = h[0] * h[1] * h[2];
  for (int i = tid; i < NS * 18; i += NT) sU[i] = U[(size_t)cn[i / 18] * 18 + (i % 18)];
  __syncthreads();
  for (int i = tid; i < NQ * 18; i += NT) {
    const int q = i / 18, c = i % 18;
    double xv = 0, g0 = 0, g1 = 0, g2 = 0;
    for (int a = 0; a < NS; ++a) {
      const double Ua = sU[a * 18 + c];
      xv += T.N[q * NS + a] * Ua;
      g0 += T.dN[(q * NS + a) * 3 + 0] * Ua;
      g1 += T.dN[(q * NS + a) * 3 + 1] * Ua;
      g2 += T.dN[(q * NS + a) * 3 + 2] * Ua;
    }
    sX[i] = xv;
    sG[q * 54 + c] = g0 * ih[0];
    sG[q * 54 + 18 + c] = g1 * ih[1];
    sG[q * 54 + 36 + c] = g2 * ih[2];
  }
  __syncthreads();
  double acc = 0.0;
  for (int q = tid; q < NQ; q += NT) {
    const double *x = sX + q * 18, *u = x, *v = x + 9, *g = sG + q * 54;
    double uut[9], vvt[9], uvt[9];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double a1 = 0, a2 = 0, a3 = 0;
        for (int k = 0; k < 3; ++k) {
          a1 += u[3 * i + k] * u[3 * j + k];
          a2 += v[3 * i + k] * v[3 * j + k];
          a3 += u[3 * i + k] * v[3 * j + k];
        }
        uut[3 * i + j] = a1; vvt[3 * i + j] = a2; uvt[3 * i + j] = a3;
      }
    double U2 = 0, V2 = 0, UV = 0;
    for (int a = 0; a < 9; ++a) { U2 += u[a] * u[a]; V2 += v[a] * v[a]; UV += u[a] * v[a]; }
    double trP2 = 0, trQ2 = 0, trS2 = 0, trW2 = 0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        const double P = uut[3 * i + j] - vvt[3 * i + j], S = uut[3 * i + j] + vvt[3 * i + j];
        const double Q = uvt[3 * i + j] + uvt[3 * j + i];
        const double Wij = uvt[3 * j + i] - uvt[3 * i + j], Wji = uvt[3 * i + j] - uvt[3 * j + i];
        trP2 += P * P; trQ2 += Q * Q; trS2 += S * S; trW2 += Wij * Wji;
      }
    const double D = U2 - V2, N = U2 + V2;
    double f = M.al * N + M.b1 * (D * D + 4 * UV * UV) + M.b2 * N * N + M.b3 * (trP2 + trQ2) +
               M.b4 * (trS2 - trW2) + M.b5 * (trS2 + trW2);
    double g2 = 0;
    for (int i = 0; i < 54; ++i) g2 += g[i] * g[i];
    double dv = 0;
    for (int part = 0; part < 2; ++part)
      for (int m = 0; m < 3; ++m) {
        const double dd = g[9 * part + 3 * m] + g[18 + 9 * part + 3 * m + 1] + g[36 + 9 * part + 3 * m + 2];
        dv += dd * dd;
      }
    f += M.K1 * g2 + M.K23 * dv;
    acc += T.w[q] * vol * f;
  }
  // Robin wall energy: (K1/bt) sum_faces U^T M_f U over unzeroed components
  if (M.robin != 0.0) {
    const uint32_t cface = Mh.cell_face[e];
    for (int lf = 0; lf < 6; ++lf) {
      const int bid = (int)((cface >> (3 * lf)) & 7u) - 1;
      if (bid < 2 || bid > 4) continue;
      for (int i = tid; i < NS * 18; i += NT) {
        const int a = i / 18, c = i % 18;
        if (face_zeroed(bid, c)) continue;
        double s = 0;
        for (int b = 0; b < NS; ++b) s += T.Mf[(lf * NS + a) * NS + b] * sU[b * 18 + c];
        acc += M.robin * face_area(lf, h) * sU[a * 18 + c] * s;
      }
    }
  }
  sF[tid] = acc;
  __syncthreads();
  // fixed-order tree: fold 192 -> 128, then halve
  if (tid < NT - 128) sF[tid] += sF[tid + 128];
  __syncthreads();
  for (int st = 64; st > 0; st >>= 1) {
    if (tid < st) sF[tid] += sF[tid + st];
    __syncthreads();
  }
  if (tid == 0) Fcell[e] = sF[0];
}

}  // namespace

hipError_t launch_assemble(int deg, bool jfp32, const RefTables &T, const MeshDev &M, const Material &mat,
                           const double *U, const int *cells, int n, void *J, hipStream_t s)
{
  if (n <= 0) return hipSuccess;
  dim3 g(n), b(NT);
#define GLH_ASM(D, TJ, HG) hipLaunchKernelGGL((k_assemble<D, TJ, HG>), g, b, 0, s, T, M, mat, U, cells, (TJ *)J)
  // the hanging-node variant (adaptive meshes) also gathers the constrained diagonals
  const bool hang = M.hdiag != nullptr;
  if (deg == 1) {
    if (hang) { if (jfp32) GLH_ASM(1, float, true); else GLH_ASM(1, double, true); }
    else { if (jfp32) GLH_ASM(1, float, false); else GLH_ASM(1, double, false); }
  } else {
    if (hang) { if (jfp32) GLH_ASM(2, float, true); else GLH_ASM(2, double, true); }
    else { if (jfp32) GLH_ASM(2, float, false); else GLH_ASM(2, double, false); }
  }
#undef GLH_ASM
  return hipGetLastError();
}

hipError_t launch_mf_setup(int deg, bool u32, const RefTables &T, const MeshDev &M, const Material &mat,
                           const double *U, const int *cells, int n, void *Uq, double *Jd, const MfCondDev &X,
                           hipStream_t s)
{
  if (n <= 0) return hipSuccess;
  if (X.xptr && !M.hdiag) return hipErrorInvalidValue;  // the constrained setup writes hdiag
#define GLH_MFS(D, TH, CON) \
  hipLaunchKernelGGL((k_mf_setup<D, TH, CON>), dim3(n), dim3(NT), 0, s, T, M, mat, U, cells, (TH *)Uq, Jd, X)
#define GLH_MFS2(D, TH) \
  if (X.xptr) GLH_MFS(D, TH, true); else GLH_MFS(D, TH, false)
  if (deg == 1) { if (u32) GLH_MFS2(1, float); else GLH_MFS2(1, double); }
  else { if (u32) GLH_MFS2(2, float); else GLH_MFS2(2, double); }
#undef GLH_MFS2
#undef GLH_MFS
  return hipGetLastError();
}

hipError_t launch_mf_apply(int deg, bool u32, bool x32, const RefTables &T, const MeshDev &M, const Material &mat,
                           const void *Uq, const void *x, const int *cells, int n, const int *flag, void *ycell,
                           hipStream_t s)
{
  if (n <= 0) return hipSuccess;
  if (x32 && !u32) return hipErrorInvalidValue;  // an FP64 operator is applied to FP64 vectors only
#define GLH_MFA(D, TH, TX)                                                                                  \
  hipLaunchKernelGGL((k_mf_apply<D, TH, TX>), dim3(n), dim3((mf_threads<D, TH>())), 0, s, T, M, mat, (const TH *)Uq, \
                     (const TX *)x, cells, flag, (mf_arith<TH, TX> *)ycell)
  if (deg == 1) {
    if (u32) { if (x32) GLH_MFA(1, float, float); else GLH_MFA(1, float, double); }
    else GLH_MFA(1, double, double);
  } else {
    if (u32) { if (x32) GLH_MFA(2, float, float); else GLH_MFA(2, float, double); }
    else GLH_MFA(2, double, double);
  }
#undef GLH_MFA
  return hipGetLastError();
}

hipError_t launch_res_apply(int deg, const RefTables &T, const MeshDev &M, const Material &mat, const double *U,
                            const int *cells, int n, double *ycell, hipStream_t s)
{
  if (n <= 0) return hipSuccess;
#define GLH_RES(D)                                                                                               \
  hipLaunchKernelGGL((k_mf_apply<D, double, double, true>), dim3(n), dim3((mf_threads<D, double>())), 0, s, T, M, \
                     mat, (const double *)nullptr, U, cells, M.inst_active, ycell)
  if (deg == 1) GLH_RES(1);
  else GLH_RES(2);
#undef GLH_RES
  return hipGetLastError();
}

hipError_t launch_res_gather(int n_nodes, const int *ptr, const int *idx, const double *ycell, const uint32_t *mask,
                             const int *row_seg, const int *active, double *rhs, hipStream_t s)
{
  if (n_nodes <= 0) return hipSuccess;
  const long long n = (long long)n_nodes * 9;
  hipLaunchKernelGGL(k_res_gather, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n_nodes, ptr, idx, ycell, mask,
                     row_seg, active, rhs);
  return hipGetLastError();
}

template <typename TX>
hipError_t launch_mf_gather(int n_nodes, const int *ptr, const int *idx, const void *ycell, const uint32_t *mask,
                            const double *Jd, const double *sh, const TX *x, const TX *b, const int *row_seg,
                            const int *flag, TX *y, hipStream_t s)
{
  if (n_nodes <= 0) return hipSuccess;
  const long long n = (long long)n_nodes * 9;
  // the slots are FP32 exactly when x is (the product's arithmetic type)
  using TY = typename std::conditional<std::is_same<TX, float>::value, float, double>::type;
  hipLaunchKernelGGL((k_mf_gather<TX, TY>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n_nodes, ptr, idx,
                     (const TY *)ycell, mask, Jd, sh, x, b, row_seg, flag, y);
  return hipGetLastError();
}
template hipError_t launch_mf_gather<double>(int, const int *, const int *, const void *, const uint32_t *,
                                             const double *, const double *, const double *, const double *,
                                             const int *, const int *, double *, hipStream_t);
template hipError_t launch_mf_gather<float>(int, const int *, const int *, const void *, const uint32_t *,
                                            const double *, const double *, const float *, const float *,
                                            const int *, const int *, float *, hipStream_t);

hipError_t launch_energy(int deg, const RefTables &T, const MeshDev &M, const Material &mat, const double *U,
                         int n_cells, double *Fcell, hipStream_t s)
{
  if (n_cells <= 0) return hipSuccess;
  if (deg == 1) hipLaunchKernelGGL(k_energy<1>, dim3(n_cells), dim3(NT), 0, s, T, M, mat, U, n_cells, Fcell);
  else hipLaunchKernelGGL(k_energy<2>, dim3(n_cells), dim3(NT), 0, s, T, M, mat, U, n_cells, Fcell);
  return hipGetLastError();
}

}  // namespace glh
