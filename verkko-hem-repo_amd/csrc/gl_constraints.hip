// Hanging-node constraints on gfx950: condensation of the cell-loop matrix
// and residual into the constrained system, and the distribution of
// constrained values from their masters.
//
// Reference semantics: AffineConstraints::distribute_local_to_global as used
// by femgl/src/assemble.cc:356-361 and residual.cc:286-289 (constrained rows
// and columns moved to the masters with the line weights; a constrained row
// keeps only its diagonal, the sum over cells of |local diagonal|), and
// AffineConstraints::distribute (solve.cc:181, iteration.cc:180).
//
// The cell loop assembles over all nodes (hanging ones included) into Jf, a
// BSR matrix over the unconstrained cell pattern; Jc = C^T Jf C is then a
// gather: every block of Jc is a weighted sum of blocks of Jf listed on the
// host.  The lists are in a fixed order, so the result is deterministic.
#include "glhip_dev.hpp"

namespace glh {
namespace {

constexpr int CT = 256;

// One thread per value of Jc.  Contribution k of block t: source block
// src[k], weight w[k], fl[k] bit 0: the row comes through a constraint line,
// bit 1: the column does.  Such contributions are dropped on the components
// of a Dirichlet-constrained master (the closed line has no such master).
template <typename TJ>
__global__ __launch_bounds__(CT) void k_condense_j(long long n_val, const int *__restrict__ gptr,
                                                   const int *__restrict__ src, const double *__restrict__ w,
                                                   const unsigned char *__restrict__ fl,
                                                   const int *__restrict__ brow, const int *__restrict__ col,
                                                   const uint32_t *__restrict__ mask, const TJ *__restrict__ Jf,
                                                   TJ *__restrict__ Jc)
{
  const long long stride = (long long)gridDim.x * CT;
  for (long long i = (long long)blockIdx.x * CT + threadIdx.x; i < n_val; i += stride) {
    const long long t = i / 324;
    const int e = (int)(i - t * 324), c = e / 18, d = e % 18;
    const bool mr = (mask[brow[t]] >> c) & 1u, ms = (mask[col[t]] >> d) & 1u;
    double acc = 0.0;
    for (int k = gptr[t]; k < gptr[t + 1]; ++k) {
      const unsigned f = fl[k];
      if (((f & 1u) && mr) || ((f & 2u) && ms)) continue;
      acc += w[k] * (double)Jf[(size_t)src[k] * 324 + e];
    }
    Jc[i] = (TJ)acc;
  }
}

// Diagonal of the hanging rows: sum over cells of |local diagonal|.
template <typename TJ>
__global__ __launch_bounds__(CT) void k_hang_diag(int n_hang, const int *__restrict__ hnode,
                                                  const int *__restrict__ dblk, const double *__restrict__ hdiag,
                                                  TJ *__restrict__ Jc)
{
  const int i = blockIdx.x * CT + threadIdx.x;
  if (i >= n_hang * 18) return;
  const int h = hnode[i / 18], c = i % 18;
  Jc[(size_t)dblk[h] * 324 + c * 19] = (TJ)hdiag[(size_t)h * 18 + c];
}

// rc[r,c] = sum_k w[k] rf[src[k], c] (k in rptr[r]..rptr[r+1]), dropping
// constraint contributions on the Dirichlet components of r
__global__ __launch_bounds__(CT) void k_condense_r(int n_nodes, const int *__restrict__ rptr,
                                                   const int *__restrict__ src, const double *__restrict__ w,
                                                   const unsigned char *__restrict__ fl,
                                                   const uint32_t *__restrict__ mask, const double *__restrict__ rf,
                                                   double *__restrict__ rc)
{
  const long long i = (long long)blockIdx.x * CT + threadIdx.x;
  if (i >= (long long)n_nodes * 18) return;
  const int r = (int)(i / 18), c = (int)(i % 18);
  const bool mr = (mask[r] >> c) & 1u;
  double acc = 0.0;
  for (int k = rptr[r]; k < rptr[r + 1]; ++k) {
    if ((fl[k] & 1u) && mr) continue;
    acc += w[k] * rf[(size_t)src[k] * 18 + c];
  }
  rc[i] = acc;
}

// v[h,c] = sum_k hw[k] v[hm[k], c] for the hanging nodes (masters are never
// hanging: the lines are closed)
__global__ __launch_bounds__(CT) void k_distribute(int n_hang, const int *__restrict__ hnode,
                                                   const int *__restrict__ hptr, const int *__restrict__ hm,
                                                   const double *__restrict__ hw, double *__restrict__ v)
{
  const int i = blockIdx.x * CT + threadIdx.x;
  if (i >= n_hang * 18) return;
  const int j = i / 18, c = i % 18;
  double acc = 0.0;
  for (int k = hptr[j]; k < hptr[j + 1]; ++k) acc += hw[k] * v[(size_t)hm[k] * 18 + c];
  v[(size_t)hnode[j] * 18 + c] = acc;
}

// ------------------------------------------------ matrix-free Jacobian
// Condensed diagonal blocks, in place: thread (t, c, d) adds to block
// tnode[t] the weighted terms of its list (in list order), each the [c][d]
// entry of a raw local block with the cell loop's Dirichlet rule applied
// (row of A or column of B constrained: 0, except |v| on the diagonal of a
// node's own (a, a) block) and the contributions through a line dropped on
// the Dirichlet components of the master (as k_condense_j drops them).
__global__ __launch_bounds__(CT) void k_mf_cond_diag(int n_targets, const int *__restrict__ tnode,
                                                     const int *__restrict__ tptr, const int *__restrict__ tslot,
                                                     const int *__restrict__ tA, const int *__restrict__ tB,
                                                     const double *__restrict__ tw,
                                                     const unsigned char *__restrict__ tfl,
                                                     const uint32_t *__restrict__ mask, const double *__restrict__ LB,
                                                     double *__restrict__ Jd)
{
  const long long i = (long long)blockIdx.x * CT + threadIdx.x;
  if (i >= (long long)n_targets * 324) return;
  const int t = (int)(i / 324), e = (int)(i % 324), c = e / 18, d = e % 18;
  const int n = tnode[t];
  const uint32_t mn = mask[n];
  double acc = 0.0;
  for (int k = tptr[t]; k < tptr[t + 1]; ++k) {
    const unsigned f = tfl[k];
    if (((f & 1u) && ((mn >> c) & 1u)) || ((f & 2u) && ((mn >> d) & 1u))) continue;
    double v = LB[(size_t)tslot[k] * 324 + ((f & 4u) ? d * 18 + c : e)];
    const bool rA = (mask[tA[k]] >> c) & 1u, cB = (mask[tB[k]] >> d) & 1u;
    if (rA || cB) v = ((f & 8u) && rA && c == d) ? fabs(v) : 0.0;
    acc += tw[k] * v;
  }
  Jd[(size_t)n * 324 + e] += acc;
}

// Block n_nodes + j of the hanging node hnode[j]: its condensed row, the
// diagonal hdiag and nothing else.
__global__ __launch_bounds__(CT) void k_mf_hang_blocks(int n_nodes, int n_hang, const int *__restrict__ hnode,
                                                       const double *__restrict__ hdiag, double *__restrict__ Jd)
{
  const long long i = (long long)blockIdx.x * CT + threadIdx.x;
  if (i >= (long long)n_hang * 324) return;
  const int j = (int)(i / 324), e = (int)(i % 324), c = e / 18, d = e % 18;
  Jd[((size_t)n_nodes + j) * 324 + e] = c == d ? hdiag[(size_t)hnode[j] * 18 + c] : 0.0;
}

// xf = C x, one thread per value.  A hanging node's value is the weighted sum
// of its masters' (the closed line), leaving out the masters' Dirichlet
// components; every other value is copied.
template <typename TV>
__global__ __launch_bounds__(CT) void k_mf_distribute(int n_nodes, const int *__restrict__ hidx,
                                                      const int *__restrict__ hptr, const int *__restrict__ hm,
                                                      const double *__restrict__ hw, const uint32_t *__restrict__ mask,
                                                      const TV *__restrict__ x, const int *__restrict__ row_seg,
                                                      const int *__restrict__ flag, TV *__restrict__ xf)
{
  const long long i = (long long)blockIdx.x * CT + threadIdx.x;
  if (i >= (long long)n_nodes * 18) return;
  const int r = (int)(i / 18), c = (int)(i % 18);
  if (flag && !flag[row_seg[r]]) return;
  const int j = hidx[r];
  if (j < 0) {
    xf[i] = x[i];
    return;
  }
  double acc = 0.0;
  for (int k = hptr[j]; k < hptr[j + 1]; ++k) {
    const int mnode = hm[k];
    if (!((mask[mnode] >> c) & 1u)) acc += hw[k] * (double)x[(size_t)mnode * 18 + c];
  }
  xf[i] = (TV)acc;
}

// y = C^T yf: an unconstrained row gathers itself and the rows of the hanging
// nodes whose lines hold it (lists of k_condense_r, the latter dropped on the
// row's Dirichlet components); a hanging row is hdiag x.  With b: y = b - that.
template <typename TV>
__global__ __launch_bounds__(CT) void k_mf_condense_y(int n_nodes, const int *__restrict__ hidx,
                                                      const int *__restrict__ rptr, const int *__restrict__ src,
                                                      const double *__restrict__ w,
                                                      const unsigned char *__restrict__ fl,
                                                      const uint32_t *__restrict__ mask,
                                                      const double *__restrict__ hdiag, const TV *__restrict__ yf,
                                                      const TV *__restrict__ x, const TV *__restrict__ b,
                                                      const int *__restrict__ row_seg, const int *__restrict__ flag,
                                                      TV *__restrict__ y)
{
  const long long i = (long long)blockIdx.x * CT + threadIdx.x;
  if (i >= (long long)n_nodes * 18) return;
  const int r = (int)(i / 18), c = (int)(i % 18);
  if (flag && !flag[row_seg[r]]) return;
  double acc;
  if (hidx[r] >= 0) {
    acc = hdiag[i] * (double)x[i];
  } else {
    const bool mr = (mask[r] >> c) & 1u;
    acc = 0.0;
    for (int k = rptr[r]; k < rptr[r + 1]; ++k) {
      if ((fl[k] & 1u) && mr) continue;
      acc += w[k] * (double)yf[(size_t)src[k] * 18 + c];
    }
  }
  y[i] = b ? (TV)((double)b[i] - acc) : (TV)acc;
}

// Continuation shift of the matrix-free Jacobian (k_diag_shift's rule on the
// diagonal blocks): sh = sigma |v| on the free components, v + sh in place.
__global__ __launch_bounds__(CT) void k_mf_shift(int n_nodes, const uint32_t *__restrict__ mask,
                                                 const int *__restrict__ row_seg, const double *__restrict__ sigma,
                                                 double *__restrict__ Jd, double *__restrict__ sh)
{
  const long long i = (long long)blockIdx.x * CT + threadIdx.x;
  if (i >= (long long)n_nodes * 18) return;
  const int r = (int)(i / 18), c = (int)(i % 18);
  const uint32_t mk = mask[r];
  if ((mk & HANG_BIT) || ((mk >> c) & 1u)) {
    sh[i] = 0.0;
    return;
  }
  double *p = Jd + (size_t)r * 324 + c * 19;
  const double v = *p, t = sigma[row_seg[r]] * fabs(v);
  sh[i] = t;
  *p = v + t;
}

// Solution transfer to a refined mesh (SolutionTransfer of FE_Q(deg),
// femgl/src/refine.cc:163-175): value (v, c) is the Q_deg interpolant of the
// old cell cell[v] at t[v], with every product and sum rounded on its own in
// the order of the host transfer, so both give the same bits.  Contraction
// into fused multiply-adds is switched off here: it would round w * U and the
// sum once (the build compiles with -ffp-contract=fast).
__device__ inline double lagrange_rn(int deg, int i, double t)
{
#pragma clang fp contract(off)
  if (deg == 1) return i ? t : 1.0 - t;
  if (i == 0) return (2.0 * (t - 0.5)) * (t - 1.0);
  if (i == 1) return (-4.0 * t) * (t - 1.0);
  return (2.0 * t) * (t - 0.5);
}

__global__ __launch_bounds__(CT) void k_transfer(int n_nodes, int deg, const int *__restrict__ cell,
                                                 const float *__restrict__ t, const int *__restrict__ old_cn,
                                                 const double *__restrict__ Uo, double *__restrict__ Un)
{
#pragma clang fp contract(off)
  const long long i = (long long)blockIdx.x * CT + threadIdx.x;
  if (i >= (long long)n_nodes * 18) return;
  const int v = (int)(i / 18), c = (int)(i % 18);
  const int n1 = deg + 1, ns = n1 * n1 * n1;
  const double t0 = t[3 * (size_t)v], t1 = t[3 * (size_t)v + 1], t2 = t[3 * (size_t)v + 2];
  const int *cn = old_cn + (size_t)cell[v] * ns;
  double u = 0.0;
  for (int a = 0; a < ns; ++a) {
    const double w = (lagrange_rn(deg, a % n1, t0) * lagrange_rn(deg, (a / n1) % n1, t1)) *
                     lagrange_rn(deg, a / (n1 * n1), t2);
    if (w != 0.0) u = u + w * Uo[(size_t)cn[a] * 18 + c];
  }
  Un[i] = u;
}

}  // namespace

hipError_t launch_transfer(int n_nodes, int deg, const int *cell, const float *t, const int *old_cn,
                           const double *Uo, double *Un, hipStream_t s)
{
  if (n_nodes <= 0) return hipSuccess;
  if (deg != 1 && deg != 2) return hipErrorInvalidValue;
  const long long n = (long long)n_nodes * 18;
  hipLaunchKernelGGL(k_transfer, dim3((unsigned)((n + CT - 1) / CT)), dim3(CT), 0, s, n_nodes, deg, cell, t, old_cn,
                     Uo, Un);
  return hipGetLastError();
}

template <typename TJ>
hipError_t launch_condense(const ConstraintDev &C, const TJ *Jf, TJ *Jc, const double *hdiag, hipStream_t s)
{
  const long long nv = (long long)C.nnzb * 324;
  long long g = (nv + CT - 1) / CT;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  hipLaunchKernelGGL((k_condense_j<TJ>), dim3((unsigned)g), dim3(CT), 0, s, nv, C.gptr, C.gsrc, C.gw, C.gfl, C.brow,
                     C.col, C.mask, Jf, Jc);
  if (C.n_hang > 0)
    hipLaunchKernelGGL((k_hang_diag<TJ>), dim3((C.n_hang * 18 + CT - 1) / CT), dim3(CT), 0, s, C.n_hang, C.hnode,
                       C.diag, hdiag, Jc);
  return hipGetLastError();
}

hipError_t launch_condense_rhs(const ConstraintDev &C, const double *rf, double *rc, hipStream_t s)
{
  const long long n = (long long)C.n_nodes * 18;
  hipLaunchKernelGGL(k_condense_r, dim3((unsigned)((n + CT - 1) / CT)), dim3(CT), 0, s, C.n_nodes, C.rptr, C.rsrc,
                     C.rw, C.rfl, C.mask, rf, rc);
  return hipGetLastError();
}

hipError_t launch_distribute(const ConstraintDev &C, double *v, hipStream_t s)
{
  if (C.n_hang <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_distribute, dim3((C.n_hang * 18 + CT - 1) / CT), dim3(CT), 0, s, C.n_hang, C.hnode, C.hptr,
                     C.hm, C.hw, v);
  return hipGetLastError();
}

hipError_t launch_mf_cond_diag(const ConstraintDev &C, const MfCondDev &X, int n_nodes, const double *hdiag,
                               double *Jd, hipStream_t s)
{
  if (X.n_targets > 0) {
    const long long nv = (long long)X.n_targets * 324;
    hipLaunchKernelGGL(k_mf_cond_diag, dim3((unsigned)((nv + CT - 1) / CT)), dim3(CT), 0, s, X.n_targets, X.tnode,
                       X.tptr, X.tslot, X.tA, X.tB, X.tw, X.tfl, C.mask, X.LB, Jd);
  }
  if (C.n_hang > 0) {
    const long long nv = (long long)C.n_hang * 324;
    hipLaunchKernelGGL(k_mf_hang_blocks, dim3((unsigned)((nv + CT - 1) / CT)), dim3(CT), 0, s, n_nodes, C.n_hang,
                       C.hnode, hdiag, Jd);
  }
  return hipGetLastError();
}

template <typename TV>
hipError_t launch_mf_distribute(const ConstraintDev &C, int n_nodes, const TV *x, TV *xf, const int *row_seg,
                                const int *flag, hipStream_t s)
{
  const long long n = (long long)n_nodes * 18;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_mf_distribute<TV>), dim3((unsigned)((n + CT - 1) / CT)), dim3(CT), 0, s, n_nodes, C.hidx,
                     C.hptr, C.hm, C.hw, C.mask, x, row_seg, flag, xf);
  return hipGetLastError();
}

template <typename TV>
hipError_t launch_mf_condense(const ConstraintDev &C, int n_nodes, const double *hdiag, const TV *yf, const TV *x,
                              const TV *b, TV *y, const int *row_seg, const int *flag, hipStream_t s)
{
  const long long n = (long long)n_nodes * 18;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_mf_condense_y<TV>), dim3((unsigned)((n + CT - 1) / CT)), dim3(CT), 0, s, n_nodes, C.hidx,
                     C.rptr, C.rsrc, C.rw, C.rfl, C.mask, hdiag, yf, x, b, row_seg, flag, y);
  return hipGetLastError();
}

hipError_t launch_mf_shift(int n_nodes, const uint32_t *mask, const int *row_seg, const double *sigma, double *Jd,
                           double *sh, hipStream_t s)
{
  const long long n = (long long)n_nodes * 18;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_mf_shift, dim3((unsigned)((n + CT - 1) / CT)), dim3(CT), 0, s, n_nodes, mask, row_seg, sigma,
                     Jd, sh);
  return hipGetLastError();
}

template hipError_t launch_mf_distribute<double>(const ConstraintDev &, int, const double *, double *, const int *,
                                                 const int *, hipStream_t);
template hipError_t launch_mf_distribute<float>(const ConstraintDev &, int, const float *, float *, const int *,
                                                const int *, hipStream_t);
template hipError_t launch_mf_condense<double>(const ConstraintDev &, int, const double *, const double *,
                                               const double *, const double *, double *, const int *, const int *,
                                               hipStream_t);
template hipError_t launch_mf_condense<float>(const ConstraintDev &, int, const double *, const float *,
                                              const float *, const float *, float *, const int *, const int *,
                                              hipStream_t);
template hipError_t launch_condense<double>(const ConstraintDev &, const double *, double *, const double *,
                                            hipStream_t);
template hipError_t launch_condense<float>(const ConstraintDev &, const float *, float *, const double *,
                                           hipStream_t);

}  // namespace glh
