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

template hipError_t launch_condense<double>(const ConstraintDev &, const double *, double *, const double *,
                                            hipStream_t);
template hipError_t launch_condense<float>(const ConstraintDev &, const float *, float *, const double *,
                                           hipStream_t);

}  // namespace glh
