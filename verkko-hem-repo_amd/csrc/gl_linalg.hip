// Linear-algebra kernels of the Newton-GMRES solve on gfx950: BSR(18x18)
// SpMV, 18x18 point-block Jacobi (inverse and apply), deterministic
// multi-dot reductions and the vector updates of GMRES and the line search.
// They replace the Trilinos Epetra SpMV / dot products inside deal.II's
// SolverFGMRES (femgl/src/solve.cc:159-174) and the vector updates of
// femgl/src/iteration.cc:170-206.
//
// Layout: vectors are node-major (dof = node*18 + c); the matrix is BSR with
// row-major 18x18 blocks of 324 contiguous values.  All kernels are
// HBM-bound; the SpMV streams each block once, every load instruction of a
// wavefront covering one contiguous span of the block.
#include "glhip_dev.hpp"

#include <mutex>
#include <vector>

namespace glh {

// Scratch of the no-allocation residual: the per-chunk partial sums of its
// norm reduction (one value per reduction chunk).  Every call on one device
// shares this array, so calls that overlap on a device race.
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

hipError_t scratch_address(int device, double **out)
{
  static std::mutex mu;
  static std::vector<double *> addr;  // per device, found once
  std::lock_guard<std::mutex> lock(mu);
  if ((int)addr.size() <= device) addr.resize(device + 1, nullptr);
  if (!addr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    addr[device] = static_cast<double *>(p);
  }
  *out = addr[device];
  return hipSuccess;
}

namespace {

template <typename T> struct V2t;
template <> struct V2t<double> { using type = double2; };
template <> struct V2t<float> { using type = float2; };

// Pair of values at a 16-byte (double) / 8-byte (float) boundary.  Matrix
// values are read once per product with non-temporal loads (streamed past the
// caches, which keep x); x is read with ordinary loads.
template <typename T> struct NV2;
template <> struct NV2<double> { typedef double type __attribute__((ext_vector_type(2))); };
template <> struct NV2<float> { typedef float type __attribute__((ext_vector_type(2))); };

template <typename T>
__device__ inline void pair_nt(const T *__restrict__ p, double &a, double &b)
{
  const typename NV2<T>::type v = __builtin_nontemporal_load(reinterpret_cast<const typename NV2<T>::type *>(p));
  a = v.x;
  b = v.y;
}

template <typename T>
__device__ inline void pair_ld(const T *__restrict__ p, double &a, double &b)
{
  const typename V2t<T>::type v = *reinterpret_cast<const typename V2t<T>::type *>(p);
  a = v.x;
  b = v.y;
}

// Block-row layout of the SpMV and the block-Jacobi apply: one wavefront per
// block row.  The 324 row-major values of a block are read as 162 consecutive
// pairs; lane l < 54 takes pairs k*54 + l (k < 3), i.e. row 6k + l/9 at
// columns 2(l%9) and 2(l%9)+1.  Each load instruction of the wave so covers 54
// consecutive pairs (864 B of FP64), and a lane needs one pair of x per block.
// The lane's three partial row sums are combined through LDS at the end of
// the row: row r = 6k + m is the sum over the 9 lanes 9m..9m+8 of their
// k-th partial sum, in lane order.
struct RowAcc {
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  template <typename TJ>
  __device__ inline void add_block(const TJ *__restrict__ blk, int lane, double w0, double w1)
  {
    double p[6];
    pair_nt(blk + 2 * lane, p[0], p[1]);
    pair_nt(blk + 2 * (54 + lane), p[2], p[3]);
    pair_nt(blk + 2 * (108 + lane), p[4], p[5]);
    a0 += p[0] * w0;
    a0 += p[1] * w1;
    a1 += p[2] * w0;
    a1 += p[3] * w1;
    a2 += p[4] * w0;
    a2 += p[5] * w1;
  }
  // lanes < 18 return the sum of row `lane`; red holds 162 doubles of this wave
  __device__ inline double row_sum(int lane, double *red) const
  {
    if (lane < 54) {
      red[lane] = a0;
      red[54 + lane] = a1;
      red[108 + lane] = a2;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    double t = 0.0;
    if (lane < 18) {
      const double *q = red + (lane / 6) * 54 + 9 * (lane % 6);
#pragma unroll
      for (int j = 0; j < 9; ++j) t += q[j];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    return t;
  }
};

__device__ inline int work_count(const ActiveSet &a, int n_chunks) { return a.counts ? a.counts[1] : n_chunks; }
__device__ inline int work_chunk(const ActiveSet &a, int i) { return a.counts ? a.chunk[i] : i; }

// block row of the g-th active row: binary search in the row prefix of the
// active instances
__device__ inline int active_row(const ActiveSet &a, int g)
{
  int lo = 0, hi = a.counts[0];
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (a.rpref[mid] <= g) lo = mid;
    else hi = mid;
  }
  return a.node_off[a.seg[lo]] + g - a.rpref[lo];
}

// Workgroups are dispatched round-robin over the 8 XCDs of the device
// (workgroup b runs on XCD b % 8, each XCD with its own L2).  The SpMV
// renumbers them so that every XCD takes one contiguous range of block rows:
// neighbouring rows share most of their x entries, which then meet in the
// same L2 instead of being fetched into all eight.
__device__ inline int xcd_block(int b, int n_blocks)
{
  const int q = n_blocks >> 3, r = n_blocks & 7, x = b & 7;
  return x * q + min(x, r) + (b >> 3);
}

// Sum of row `lane` (lanes < 18) of block row `row` times x; two blocks per
// step keep six pair loads of the matrix in flight per lane.
template <typename TJ, typename TX>
__device__ inline double spmv_row(int row, int lane, const int *__restrict__ rp, const int *__restrict__ col,
                                  const TJ *__restrict__ val, const TX *__restrict__ x, double *red)
{
  RowAcc acc;
  const int b0 = rp[row], b1 = rp[row + 1];
  if (lane < 54) {
    const int cp = 2 * (lane % 9);
    int b = b0;
    for (; b + 1 < b1; b += 2) {
      const int s0 = col[b], s1 = col[b + 1];
      double w0, w1, z0, z1;
      pair_ld(x + (size_t)s0 * 18 + cp, w0, w1);
      pair_ld(x + (size_t)s1 * 18 + cp, z0, z1);
      acc.add_block(val + (size_t)b * 324, lane, w0, w1);
      acc.add_block(val + (size_t)(b + 1) * 324, lane, z0, z1);
    }
    if (b < b1) {
      double w0, w1;
      pair_ld(x + (size_t)col[b] * 18 + cp, w0, w1);
      acc.add_block(val + (size_t)b * 324, lane, w0, w1);
    }
  }
  return acc.row_sum(lane, red);
}

// Without an active set: one row per wave, grid = rows / 4.  With one:
// grid-stride over the block rows of the listed instances.  Rows of
// instances whose seg_flag is 0 are skipped.
template <typename TJ, typename TX>
__global__ __launch_bounds__(256) void k_spmv(int n_rows, const int *__restrict__ rp, const int *__restrict__ col,
                                              const TJ *__restrict__ val, const TX *__restrict__ x,
                                              TX *__restrict__ y, const int *__restrict__ row_seg,
                                              const int *__restrict__ seg_flag, ActiveSet act)
{
  __shared__ double red[4][162];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (!act.counts) {
    const int row = xcd_block(blockIdx.x, gridDim.x) * 4 + wv;
    if (row >= n_rows) return;
    if (seg_flag && !seg_flag[row_seg[row]]) return;  // instance not iterating
    const double t = spmv_row(row, lane, rp, col, val, x, red[wv]);
    if (lane < 18) y[(size_t)row * 18 + lane] = (TX)t;
    return;
  }
  const int total = act.counts[2];
  for (int gi = blockIdx.x * 4 + wv; gi < total; gi += gridDim.x * 4) {
    const int row = active_row(act, gi);
    if (seg_flag && !seg_flag[row_seg[row]]) continue;
    const double t = spmv_row(row, lane, rp, col, val, x, red[wv]);
    if (lane < 18) y[(size_t)row * 18 + lane] = (TX)t;
  }
}

// r = b - J x with FP64 vectors (used for the outer iterative refinement with
// an FP32 matrix and for the GMRES restart residual)
template <typename TJ>
__global__ __launch_bounds__(256) void k_res_spmv(int n_rows, const int *__restrict__ rp, const int *__restrict__ col,
                                                  const TJ *__restrict__ val, const double *__restrict__ x,
                                                  const double *__restrict__ bv, double *__restrict__ r)
{
  __shared__ double red[4][162];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int row = xcd_block(blockIdx.x, gridDim.x) * 4 + wv;
  if (row >= n_rows) return;
  const double t = spmv_row(row, lane, rp, col, val, x, red[wv]);
  if (lane < 18) r[(size_t)row * 18 + lane] = bv[(size_t)row * 18 + lane] - t;
}

// Order the LDS accesses of one wavefront around a point where its lanes
// exchange values (the wavefront runs in step; no workgroup barrier).
__device__ inline void wave_lds_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Gauss-Jordan inverse of the 18x18 diagonal blocks with partial pivoting,
// in place: three nodes per wavefront, lane j of a node's 18 holding column j
// of its block in registers.  A pivot step is one division and 17 FMAs per
// lane; the pivot row and value come from the pivot column's lane by
// shuffles, that column's multipliers through LDS, and no workgroup barrier
// is needed.  The row interchanges are undone as column interchanges at the
// end.  The in-place form keeps exactly the right half of the augmented
// [A | I] elimination (same pivots, same divisions and FMAs in the same
// order), so the inverse is that elimination's to the bit.
template <typename TJ>
__global__ __launch_bounds__(256) void k_bj_invert(int n_rows, const int *__restrict__ diag,
                                                   const TJ *__restrict__ val, TJ *__restrict__ Dinv,
                                                   int *__restrict__ status)
{
  __shared__ double sC[4][4][18];     // the multipliers of the current step, per node (slot 3: idle lanes)
  __shared__ double sM[4][3][18 * 18]; // the unscrambled inverse, for the symmetric packed store
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int sub = lane / 18, j = lane - 18 * sub;  // sub 3: lanes 54..63 hold no node
  const long long node = ((long long)blockIdx.x * 4 + wv) * 3 + sub;
  const bool valid = sub < 3 && node < n_rows;
  const int base = 18 * sub;
  double a[18];
  {
    const TJ *A = val + (valid ? (size_t)diag[node] * 324 : 0);
#pragma unroll
    for (int i = 0; i < 18; ++i) a[i] = valid ? (double)A[i * 18 + j] : (i == j ? 1.0 : 0.0);
  }
  int perm[18];
  bool singular = false;
#pragma unroll
  for (int c = 0; c < 18; ++c) {
    // pivot: the first row r >= c of largest |A[r][c]| (column c is lane base + c)
    int pl = c;
    double best = fabs(a[c]);
#pragma unroll
    for (int r = c + 1; r < 18; ++r)
      if (fabs(a[r]) > best) { best = fabs(a[r]); pl = r; }
    const int p = __shfl(pl, base + c);
    perm[c] = p;
#pragma unroll
    for (int r = c + 1; r < 18; ++r) {  // interchange rows c and p of this column
      const bool sw = p == r;
      const double t = a[c];
      a[c] = sw ? a[r] : t;
      a[r] = sw ? t : a[r];
    }
    const double piv = __shfl(a[c], base + c);
    singular = singular || piv == 0.0;
    if (j == c)
#pragma unroll
      for (int i = 0; i < 18; ++i) sC[wv][sub][i] = a[i];  // the multipliers A[i][c], i != c
    a[c] = (j == c ? 1.0 : a[c]) / piv;                     // row c / pivot (column c: 1 / pivot)
    wave_lds_sync();
#pragma unroll
    for (int i = 0; i < 18; ++i)
      if (i != c) a[i] = (j == c ? 0.0 : a[i]) - sC[wv][sub][i] * a[c];
    wave_lds_sync();  // this step's reads before the next step's stores
  }
  // undo the row interchanges as column interchanges, last step first
  int pos = j;
#pragma unroll
  for (int c = 17; c >= 0; --c) {
    if (pos == c) pos = perm[c];
    else if (pos == perm[c]) pos = c;
  }
  if (sub < 3) {
    double *M = sM[wv][sub];
#pragma unroll
    for (int i = 0; i < 18; ++i) M[i * 18 + pos] = a[i];
  }
  wave_lds_sync();
  if (!valid) return;
  if (singular) {
    if (j == 0) status[0] = 1;
    return;
  }
  // the inverse of a symmetric block is symmetric: its upper triangle,
  // row-major over c <= d, the two rounded halves averaged
  const double *M = sM[wv][sub];
  TJ *D = Dinv + (size_t)node * 171;
  for (int q = j; q < 171; q += 18) {
    int c = 0, r = q;
    while (r >= 18 - c) { r -= 18 - c; ++c; }
    const int d = c + r;
    D[q] = (TJ)(0.5 * (M[c * 18 + d] + M[d * 18 + c]));
  }
}

// Point-block Jacobi y = D^-1 x with the symmetric inverses stored as their
// upper triangles (171 values per node, row-major over c <= d): 53% of the
// bytes of the full blocks.  One wavefront takes three nodes; lane l < 54
// owns row c = l % 18 of node l / 18 and forms y_c = sum_d D[min][max] x_d,
// walking the packed index (steps 17 - d while d < c, then 1) in LDS, where
// the node's 171 values and x of the three nodes are staged.
template <typename TJ, typename TX>
__device__ inline void bj_rows(int row, bool valid, int lane, const TJ *__restrict__ Dinv, const TX *__restrict__ x,
                               TX *__restrict__ y, double *sx)
{
  const int sub = lane / 18, c = lane % 18;
  if (lane < 54) sx[lane] = valid ? (double)x[(size_t)row * 18 + c] : 0.0;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // the node's 171 values staged through LDS by its 18 lanes, each load
  // instruction covering 18 consecutive values of each node: the row walk
  // below touches scattered entries, which from memory cost a cache-line
  // request per lane (bit-identical; 94 -> 80 us FP64, 65 -> 39 us FP32 on
  // the 32^3 Q2 cube)
  TJ *sp = reinterpret_cast<TJ *>(sx + 54) + sub * 172;
  if (lane < 54 && valid) {
    const TJ *P = Dinv + (size_t)row * 171;
#pragma unroll
    for (int t = 0; t < 10; ++t)
      if (c + 18 * t < 171) sp[c + 18 * t] = P[c + 18 * t];
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (lane < 54 && valid) {
    const TJ *P = sp;
    const double *xs = sx + sub * 18;
    double acc = 0.0;
    int off = c;
#pragma unroll
    for (int d = 0; d < 18; ++d) {
      acc += (double)P[off] * xs[d];
      off += d < c ? 17 - d : 1;
    }
    y[(size_t)row * 18 + c] = (TX)acc;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename TJ, typename TX>
__global__ __launch_bounds__(256) void k_bj_apply(int n_rows, const TJ *__restrict__ Dinv, const TX *__restrict__ x,
                                                  TX *__restrict__ y, const int *__restrict__ row_seg,
                                                  const int *__restrict__ seg_flag, ActiveSet act)
{
  __shared__ double sx[4][54 + 3 * 172];  // x of the wave's three nodes, then their staged inverses
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, sub = lane < 54 ? lane / 18 : 0;
  auto live = [&](int row) { return !seg_flag || seg_flag[row_seg[row]]; };
  if (!act.counts) {
    const int row = (blockIdx.x * 4 + wv) * 3 + sub;
    const bool valid = lane < 54 && row < n_rows && live(row);
    bj_rows(row, valid, lane, Dinv, x, y, sx[wv]);
    return;
  }
  const int total = act.counts[2];
  for (int g = blockIdx.x * 4 + wv; 3 * g < total; g += gridDim.x * 4) {
    const int gi = 3 * g + sub;
    const bool in = lane < 54 && gi < total;
    const int row = in ? active_row(act, gi) : 0;
    bj_rows(row, in && live(row), lane, Dinv, x, y, sx[wv]);
  }
}

// -------------------------------------------------------------- reductions
constexpr int RT = 256;
constexpr int KCH = 8;  // vectors per multi-dot pass (4 measured the same, 16 slower)
// pairs of values per thread of the Gram-Schmidt update: a chunk holds at
// most GS_MAX_CHUNK values
constexpr int GS_PAIRS = GS_MAX_CHUNK / (2 * RT);
static_assert(GS_PAIRS * 2 * RT == GS_MAX_CHUNK, "chunk bound in whole pairs per thread");

// Sum over the 64 lanes of a wave (butterfly: every lane gets the same,
// order-fixed result).
template <typename T>
__device__ inline T wave_sum(T v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Sum over the RT threads of a workgroup, returned to every thread, in fixed
// order (deterministic).  sh: at least RT/64 entries.
template <typename T>
__device__ inline T block_reduce(T v, T *sh)
{
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  T r = 0;
#pragma unroll
  for (int i = 0; i < RT / 64; ++i) r += sh[i];
  __syncthreads();
  return r;
}

// partial[blk*k + i] = sum over this block's fixed index set of V_i . w
template <typename T>
__global__ __launch_bounds__(RT) void k_mdot_partial(size_t n, int k, int i0, const T *__restrict__ V, size_t ldv,
                                                     const T *__restrict__ w, T *__restrict__ partial)
{
  __shared__ T sh[RT];
  T acc[KCH];
#pragma unroll
  for (int j = 0; j < KCH; ++j) acc[j] = 0;
  const int kk = (k - i0) < KCH ? (k - i0) : KCH;
  const size_t stride = (size_t)gridDim.x * RT;
  for (size_t idx = (size_t)blockIdx.x * RT + threadIdx.x; idx < n; idx += stride) {
    const T wv = w[idx];
#pragma unroll
    for (int j = 0; j < KCH; ++j)
      if (j < kk) acc[j] += V[(size_t)(i0 + j) * ldv + idx] * wv;
  }
  for (int j = 0; j < kk; ++j) {
    const T r = block_reduce(acc[j], sh);
    if (threadIdx.x == 0) partial[(size_t)blockIdx.x * k + i0 + j] = r;
  }
}

// out[i] (+)= sum over blocks of partial[blk*k + i], fixed order
template <typename T>
__global__ __launch_bounds__(RT) void k_mdot_final(int nblk, int k, const T *__restrict__ partial, T *__restrict__ out,
                                                   int accumulate)
{
  __shared__ T sh[RT];
  const int i = blockIdx.x;
  T acc = 0;
  for (int b = threadIdx.x; b < nblk; b += RT) acc += partial[(size_t)b * k + i];
  const T r = block_reduce(acc, sh);
  if (threadIdx.x == 0) out[i] = accumulate ? out[i] + r : r;
}

template <typename T>
__global__ __launch_bounds__(RT) void k_sub_lincomb(size_t n, int k, const T *__restrict__ V, size_t ldv,
                                                    const T *__restrict__ h, T *__restrict__ w)
{
  const size_t stride = (size_t)gridDim.x * RT;
  for (size_t idx = (size_t)blockIdx.x * RT + threadIdx.x; idx < n; idx += stride) {
    T s = w[idx];
    for (int i = 0; i < k; ++i) s -= h[i] * V[(size_t)i * ldv + idx];
    w[idx] = s;
  }
}

// w[j] -= sum_i h[i] V[i*ldv + j], and the block partial of ||w||^2 after the
// update (fixed index sets: deterministic)
template <typename T>
__global__ __launch_bounds__(RT) void k_sub_lincomb_norm(size_t n, int k, const T *__restrict__ V, size_t ldv,
                                                         const T *__restrict__ h, T *__restrict__ w,
                                                         T *__restrict__ partial)
{
  __shared__ T sh[RT];
  T acc = 0;
  const size_t stride = (size_t)gridDim.x * RT;
  for (size_t idx = (size_t)blockIdx.x * RT + threadIdx.x; idx < n; idx += stride) {
    T s = w[idx];
    for (int i = 0; i < k; ++i) s -= h[i] * V[(size_t)i * ldv + idx];
    w[idx] = s;
    acc += s * s;
  }
  const T r = block_reduce(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

template <typename T>
__global__ __launch_bounds__(RT) void k_lincomb(size_t n, int k, const T *__restrict__ V, size_t ldv,
                                                const T *__restrict__ y, T *__restrict__ w)
{
  const size_t stride = (size_t)gridDim.x * RT;
  for (size_t idx = (size_t)blockIdx.x * RT + threadIdx.x; idx < n; idx += stride) {
    T s = 0;
    for (int i = 0; i < k; ++i) s += y[i] * V[(size_t)i * ldv + idx];
    w[idx] = s;
  }
}

template <typename T>
__global__ __launch_bounds__(RT) void k_scale_by(size_t n, const T *__restrict__ src, const T *__restrict__ scal,
                                                 T *__restrict__ dst)
{
  const T a = scal[0];
  const T inv = a != (T)0 ? (T)1 / a : (T)1;
  const size_t stride = (size_t)gridDim.x * RT;
  for (size_t idx = (size_t)blockIdx.x * RT + threadIdx.x; idx < n; idx += stride) dst[idx] = src[idx] * inv;
}

template <typename TX, typename TY>
__global__ __launch_bounds__(RT) void k_axpy(size_t n, double a, const TX *__restrict__ x, TY *__restrict__ y)
{
  const size_t stride = (size_t)gridDim.x * RT;
  for (size_t idx = (size_t)blockIdx.x * RT + threadIdx.x; idx < n; idx += stride)
    y[idx] = (TY)((double)y[idx] + a * (double)x[idx]);
}

template <typename TS, typename TD>
__global__ __launch_bounds__(RT) void k_convert(size_t n, const TS *__restrict__ s, TD *__restrict__ d)
{
  const size_t stride = (size_t)gridDim.x * RT;
  for (size_t idx = (size_t)blockIdx.x * RT + threadIdx.x; idx < n; idx += stride) d[idx] = (TD)s[idx];
}

template <typename TJ>
__global__ __launch_bounds__(RT) void k_gather_blocks(int n_rows, const int *__restrict__ blk,
                                                      const TJ *__restrict__ val, double *__restrict__ dst)
{
  const long long i = (long long)blockIdx.x * RT + threadIdx.x;
  if (i >= (long long)n_rows * 324) return;
  const long long r = i / 324;
  dst[i] = (double)val[(size_t)blk[r] * 324 + (i - r * 324)];
}

__global__ __launch_bounds__(RT) void k_sub(size_t n, const double *__restrict__ b, const double *__restrict__ w,
                                           double *__restrict__ d)
{
  const size_t stride = (size_t)gridDim.x * RT;
  for (size_t idx = (size_t)blockIdx.x * RT + threadIdx.x; idx < n; idx += stride) d[idx] = b[idx] - w[idx];
}

__global__ __launch_bounds__(RT) void k_trial(size_t n, const double *__restrict__ U, const double *__restrict__ dU,
                                             double a, const uint32_t *__restrict__ mask, double *__restrict__ X)
{
  const size_t stride = (size_t)gridDim.x * RT;
  for (size_t idx = (size_t)blockIdx.x * RT + threadIdx.x; idx < n; idx += stride) {
    const uint32_t m = mask[idx / 18];
    X[idx] = ((m >> (idx % 18)) & 1u) ? 0.0 : U[idx] + a * dU[idx];
  }
}

__global__ __launch_bounds__(RT) void k_mask_zero(size_t n, const uint32_t *__restrict__ mask, double *__restrict__ x)
{
  const size_t stride = (size_t)gridDim.x * RT;
  for (size_t idx = (size_t)blockIdx.x * RT + threadIdx.x; idx < n; idx += stride)
    if ((mask[idx / 18] >> (idx % 18)) & 1u) x[idx] = 0.0;
}

__global__ __launch_bounds__(RT) void k_sum_partial(size_t n, const double *__restrict__ v,
                                                   double *__restrict__ partial)
{
  __shared__ double sh[RT];
  double acc = 0.0;
  const size_t stride = (size_t)gridDim.x * RT;
  for (size_t idx = (size_t)blockIdx.x * RT + threadIdx.x; idx < n; idx += stride) acc += v[idx];
  const double r = block_reduce(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

__global__ __launch_bounds__(RT) void k_copy16(size_t n16, const float4 *__restrict__ src, float4 *__restrict__ dst)
{
  const size_t stride = (size_t)gridDim.x * RT;
  for (size_t idx = (size_t)blockIdx.x * RT + threadIdx.x; idx < n16; idx += stride) dst[idx] = src[idx];
}

// streaming read of n16 16-byte words: every workgroup reads contiguous
// blocks of RT * RD words, each thread RD independent non-temporal loads in
// flight (the fastest form measured on the device: 7.0 TB/s with 1024
// workgroups, against 5.9 TB/s with ordinary loads); one float per thread is
// written so that the loads stay live
constexpr int RD = 8;
__global__ __launch_bounds__(RT) void k_read16(size_t n16, const float4 *__restrict__ src, float *__restrict__ out)
{
  typedef float f4 __attribute__((ext_vector_type(4)));
  const f4 *in = reinterpret_cast<const f4 *>(src);
  float acc = 0.f;
  const size_t blk = (size_t)RT * RD, nblk = (n16 + blk - 1) / blk;
  for (size_t b = blockIdx.x; b < nblk; b += gridDim.x) {
    const size_t base = b * blk + threadIdx.x;
    f4 v[RD];
#pragma unroll
    for (int r = 0; r < RD; ++r) {
      const size_t idx = base + (size_t)r * RT;
      v[r] = idx < n16 ? __builtin_nontemporal_load(in + idx) : f4{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int r = 0; r < RD; ++r) acc += (v[r].x + v[r].y) + (v[r].z + v[r].w);
  }
  out[(size_t)blockIdx.x * RT + threadIdx.x] = acc;
}

template <typename T>
__global__ void k_set_scalar(T *p, T v)
{
  if (threadIdx.x == 0) p[0] = v;
}


// ------------------------------------------------ segmented (batched) kernels
// The dof range is cut into chunks that never straddle an instance boundary:
// ch[3i] = instance, ch[3i+1] = first dof, ch[3i+2] = one past the last dof.
// A workgroup handles one chunk at a time, either every chunk (grid =
// n_chunks) or, with an active set, the chunks of the instances listed in it
// (grid-stride over the compacted list, so skipped instances cost no
// workgroups).  Per-chunk partials are reduced per instance in fixed order
// (deterministic).

// Chunk bounds are even (an instance holds 18 values per node and chunks are
// 1024 values long), so the Gram-Schmidt kernels below read pairs of values:
// the basis vectors with non-temporal loads (a cycle's basis is far larger
// than the caches and is streamed), w with ordinary loads (every group of
// basis vectors reads it again).
template <typename T>
__device__ inline typename NV2<T>::type ld2nt(const T *__restrict__ p)
{
  return __builtin_nontemporal_load(reinterpret_cast<const typename NV2<T>::type *>(p));
}
template <typename T>
__device__ inline typename NV2<T>::type ld2(const T *__restrict__ p)
{
  return *reinterpret_cast<const typename NV2<T>::type *>(p);
}

// 2D grid: x = chunks (or the active list), y = groups of G vectors.
// Without an active set the grid is numbered XCD by XCD: workgroups go to
// the 8 XCDs in turn (linear id mod 8), so linear id L = 8 r + x runs chunk
// 8 (r / groups) + x, group r % groups, and every group of a chunk runs on
// the XCD of that chunk, whose L2 then serves w to all but the first
// (gridDim.x is a multiple of 8 there).
template <typename T, int G>
__global__ __launch_bounds__(RT) void k_seg_mdot_partial(const long long *__restrict__ ch, int n_chunks,
                                                         ActiveSet act, int k, const T *__restrict__ V, size_t ldv,
                                                         const T *__restrict__ w, T *__restrict__ partial,
                                                         const int *__restrict__ flag, Op1In<T> O)
{
  __shared__ T sh[G][RT / 64];
  __shared__ T shr[RT];
  __shared__ int s_cy;
  int o_sg = -1;  // O.npart: the instance whose flag this workgroup formed last
  bool o_on = false;
  int nw = work_count(act, n_chunks), first = blockIdx.x, step = gridDim.x, grp = blockIdx.y;
  if (!act.counts && (int)gridDim.x >= nw) {  // (a capped grid strides over the chunks instead)
    const int L = blockIdx.x + gridDim.x * blockIdx.y, r = L >> 3;
    grp = r % gridDim.y;
    first = 8 * (r / gridDim.y) + (L & 7);
    step = nw;  // one chunk per workgroup
  }
  const int i0 = grp * G;
  const int kk = (k - i0) < G ? (k - i0) : G;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int wi = first; wi < nw; wi += step) {
    const int cb = work_chunk(act, wi);
    if (O.npart) {
      // control op 1 of the instance here (folded batches, the second pass):
      // its norm after the first pass from that pass's chunk partials, with
      // k_seg_final_ctl's arithmetic, then the pass's flag; the instance's
      // first chunk (vector group 0) records both for the later kernels
      const int sg = (int)ch[3 * cb];
      if (sg != o_sg) {
        if (threadIdx.x == 0) s_cy = O.C.cyc[sg];
        __syncthreads();
        const bool cy = s_cy != 0;
        __syncthreads();
        T nrm = 0;
        if (cy) {
          T a = 0;
          for (int b = O.seg_chunk[sg] + threadIdx.x; b < O.seg_chunk[sg + 1]; b += RT) a += O.npart[b];
          nrm = block_reduce(a, shr);
        }
        o_on = cy && (double)nrm < O.C.dgks * (double)O.C.hs[(size_t)sg * O.C.ldh + O.j + 1];
        o_sg = sg;
        if (cb == O.seg_chunk[sg] && grp == 0 && threadIdx.x == 0) {
          if (cy) O.norm2[sg] = nrm;
          O.C.ro[sg] = o_on ? 1 : 0;
        }
      }
      if (!o_on) continue;
    } else if (flag && !flag[ch[3 * cb]]) {
      continue;  // skipped instance: its partials are never read
    }
    const long long beg = ch[3 * cb + 1], end = ch[3 * cb + 2];
    T acc[G];
#pragma unroll
    for (int j = 0; j < G; ++j) acc[j] = 0;
    for (long long idx = beg + 2 * threadIdx.x; idx < end; idx += 2 * RT) {
      const auto x = ld2(w + idx);
#pragma unroll
      for (int j = 0; j < G; ++j)
        if (j < kk) {
          const auto v = ld2nt(V + (size_t)(i0 + j) * ldv + idx);
          acc[j] += v.x * x.x;
          acc[j] += v.y * x.y;
        }
    }
#pragma unroll
    for (int j = 0; j < G; ++j) {
      const T r = wave_sum(acc[j]);
      if (lane == 0) sh[j][wv] = r;
    }
    __syncthreads();
    if ((int)threadIdx.x < kk) {
      T r = 0;
#pragma unroll
      for (int q = 0; q < RT / 64; ++q) r += sh[threadIdx.x][q];
      partial[(size_t)cb * k + i0 + threadIdx.x] = r;
    }
    __syncthreads();
  }
}

// out[s*ldo + off + i] (+)= sum over the chunks of instance s of partial[chunk*k + i]
template <typename T>
__global__ __launch_bounds__(RT) void k_seg_final(const int *__restrict__ seg_chunk, int n_seg, ActiveSet act, int k,
                                                  const T *__restrict__ partial, T *__restrict__ out, int ldo, int off,
                                                  int accumulate, const int *__restrict__ flag)
{
  __shared__ T sh[RT];
  const int i = blockIdx.y;
  const int na = act.counts ? act.counts[0] : n_seg;
  for (int a = blockIdx.x; a < na; a += gridDim.x) {
    const int sg = act.counts ? act.seg[a] : a;
    if (flag && !flag[sg]) continue;
    T acc = 0;
    for (int b = seg_chunk[sg] + threadIdx.x; b < seg_chunk[sg + 1]; b += RT) acc += partial[(size_t)b * k + i];
    const T r = block_reduce(acc, sh);
    if (threadIdx.x == 0) {
      T *o = out + (size_t)sg * ldo + off + i;
      *o = accumulate ? *o + r : r;
    }
  }
}

template <typename T, int UF>
__global__ __launch_bounds__(RT) void k_seg_sub_lincomb_norm(const long long *__restrict__ ch, int n_chunks,
                                                             ActiveSet act, int k, const T *__restrict__ V,
                                                             size_t ldv, T *h, int ldh, T *__restrict__ w,
                                                             T *__restrict__ partial, const int *__restrict__ flag,
                                                             const int *__restrict__ seg_chunk,
                                                             const T *__restrict__ hpart, int hk)
{
  __shared__ T sh[RT];
  __shared__ T shh[128];
  const int nw = work_count(act, n_chunks);
  for (int wi = blockIdx.x; wi < nw; wi += gridDim.x) {
    const int cb = work_chunk(act, wi);
    const long long sg = ch[3 * cb], beg = ch[3 * cb + 1], end = ch[3 * cb + 2];
    if (flag && !flag[sg]) continue;
    const T *hs = h + sg * ldh;
    if (hpart) {
      // the instance's hk multi-dot values from its chunk partials, in chunk
      // order (every workgroup of the instance forms the same sums); the
      // instance's first chunk stores them for the control ops
      const int c0 = seg_chunk[sg], c1 = seg_chunk[sg + 1];
      for (int i = threadIdx.x; i < hk; i += RT) {
        T acc = 0;
        for (int b = c0; b < c1; ++b) acc += hpart[(size_t)b * hk + i];
        shh[i] = acc;
        if (cb == c0) h[sg * ldh + i] = acc;
      }
      __syncthreads();
      hs = shh;
    }
    T acc = 0;
    // the updated pairs are stored once the whole chunk is done: a store
    // between the loads of the next pair holds them back (loads and stores
    // share one counter), and the stores then go out in one burst
    typename NV2<T>::type out[GS_PAIRS];
#pragma unroll
    for (int it = 0; it < GS_PAIRS; ++it) {
      const long long idx = beg + 2 * threadIdx.x + (long long)it * 2 * RT;
      if (idx >= end) break;
      const auto w2 = ld2(w + idx);
      T s0 = w2.x, s1 = w2.y;
      const T *vp = V + idx;
      int i = 0;
      for (; i + UF <= k; i += UF) {  // UF basis vectors in flight
        typename NV2<T>::type v[UF];
#pragma unroll
        for (int u = 0; u < UF; ++u) v[u] = ld2nt(vp + (size_t)(i + u) * ldv);
#pragma unroll
        for (int u = 0; u < UF; ++u) {
          s0 -= hs[i + u] * v[u].x;
          s1 -= hs[i + u] * v[u].y;
        }
      }
      for (; i < k; ++i) {
        const auto v = ld2nt(vp + (size_t)i * ldv);
        s0 -= hs[i] * v.x;
        s1 -= hs[i] * v.y;
      }
      out[it].x = s0;
      out[it].y = s1;
      acc += s0 * s0;
      acc += s1 * s1;
    }
#pragma unroll
    for (int it = 0; it < GS_PAIRS; ++it) {
      const long long idx = beg + 2 * threadIdx.x + (long long)it * 2 * RT;
      if (idx < end) *reinterpret_cast<typename NV2<T>::type *>(w + idx) = out[it];
    }
    const T r = block_reduce(acc, sh);
    if (threadIdx.x == 0) partial[cb] = r;
  }
}

template <typename T>
__global__ __launch_bounds__(RT) void k_seg_lincomb(const long long *__restrict__ ch, int n_chunks, ActiveSet act,
                                                    int k, const T *__restrict__ V, size_t ldv,
                                                    const T *__restrict__ y, int ldy, T *__restrict__ w,
                                                    const int *__restrict__ kseg)
{
  const int nw = work_count(act, n_chunks);
  for (int wi = blockIdx.x; wi < nw; wi += gridDim.x) {
    const int cb = work_chunk(act, wi);
    const long long sg = ch[3 * cb], beg = ch[3 * cb + 1], end = ch[3 * cb + 2];
    const T *ys = y + sg * ldy;
    const int kk = kseg ? (kseg[sg] < k ? kseg[sg] : k) : k;  // basis vectors this instance built
    for (long long idx = beg + threadIdx.x; idx < end; idx += RT) {
      T s = 0;
      for (int i = 0; i < kk; ++i) s += ys[i] * V[(size_t)i * ldv + idx];
      w[idx] = s;
    }
  }
}

// dst = src / scal[s*lds] (unchanged when the scale is 0); with sq != 0 the
// scale is sqrt(scal)
template <typename T>
__global__ __launch_bounds__(RT) void k_seg_scale(const long long *__restrict__ ch, int n_chunks, ActiveSet act,
                                                  const T *__restrict__ src, const T *__restrict__ scal, int lds,
                                                  T *__restrict__ dst, const int *__restrict__ flag, int sq)
{
  const int nw = work_count(act, n_chunks);
  for (int wi = blockIdx.x; wi < nw; wi += gridDim.x) {
    const int cb = work_chunk(act, wi);
    const long long sg = ch[3 * cb], beg = ch[3 * cb + 1], end = ch[3 * cb + 2];
    if (flag && !flag[sg]) continue;
    T a = scal[sg * lds];
    if (sq) a = a > (T)0 ? sqrt(a) : (T)0;
    const T inv = a != (T)0 ? (T)1 / a : (T)1;
    for (long long idx = beg + threadIdx.x; idx < end; idx += RT) dst[idx] = src[idx] * inv;
  }
}

// y += x on the (active) chunks
template <typename T>
__global__ __launch_bounds__(RT) void k_seg_add(const long long *__restrict__ ch, int n_chunks, ActiveSet act,
                                                const T *__restrict__ x, T *__restrict__ y)
{
  const int nw = work_count(act, n_chunks);
  for (int wi = blockIdx.x; wi < nw; wi += gridDim.x) {
    const int cb = work_chunk(act, wi);
    const long long beg = ch[3 * cb + 1], end = ch[3 * cb + 2];
    for (long long idx = beg + threadIdx.x; idx < end; idx += RT) y[idx] += x[idx];
  }
}

// X = U + alpha[s] dU with Dirichlet components zeroed
__global__ __launch_bounds__(RT) void k_seg_trial(const long long *__restrict__ ch, const double *__restrict__ U,
                                                 const double *__restrict__ dU, const double *__restrict__ alpha,
                                                 const uint32_t *__restrict__ mask, double *__restrict__ X)
{
  const long long sg = ch[3 * blockIdx.x], beg = ch[3 * blockIdx.x + 1], end = ch[3 * blockIdx.x + 2];
  const double a = alpha[sg];
  for (long long idx = beg + threadIdx.x; idx < end; idx += RT) {
    const uint32_t m = mask[idx / 18];
    X[idx] = ((m >> (idx % 18)) & 1u) ? 0.0 : U[idx] + a * dU[idx];
  }
}

// v = 0 on the instances whose flag is 0
template <typename T>
__global__ __launch_bounds__(RT) void k_seg_zero(const long long *__restrict__ ch, const int *__restrict__ flag,
                                                 T *__restrict__ v)
{
  const long long sg = ch[3 * blockIdx.x], beg = ch[3 * blockIdx.x + 1], end = ch[3 * blockIdx.x + 2];
  if (flag[sg]) return;
  for (long long idx = beg + threadIdx.x; idx < end; idx += RT) v[idx] = (T)0;
}

// Compacted work lists of the instances whose flag is set: one workgroup of
// AT threads scans the instances (ids and block-row prefix) and the chunks.
constexpr int AT = 1024;
__device__ inline int block_scan_incl(int v, int *sh)
{
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < AT; off <<= 1) {
    const int u = t >= off ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += u;
    __syncthreads();
  }
  return sh[t];
}

__global__ __launch_bounds__(AT) void k_build_active(int n_seg, int n_chunks, const long long *__restrict__ ch,
                                                     const int *__restrict__ flag, const int *__restrict__ node_off,
                                                     int *__restrict__ counts, int *__restrict__ seg,
                                                     int *__restrict__ rpref, int *__restrict__ chunk)
{
  __shared__ int sa[AT], sb[AT];
  int base_n = 0, base_r = 0;
  for (int t0 = 0; t0 < n_seg; t0 += AT) {
    const int sgi = t0 + threadIdx.x;
    const int f = (sgi < n_seg && flag[sgi]) ? 1 : 0;
    const int r = f ? node_off[sgi + 1] - node_off[sgi] : 0;
    const int pa = block_scan_incl(f, sa);
    const int pb = block_scan_incl(r, sb);
    if (f) {
      seg[base_n + pa - 1] = sgi;
      rpref[base_n + pa - 1] = base_r + pb - r;
    }
    base_n += sa[AT - 1];
    base_r += sb[AT - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    counts[0] = base_n;
    counts[2] = base_r;
    rpref[base_n] = base_r;
  }
  int base_c = 0;
  for (int t0 = 0; t0 < n_chunks; t0 += AT) {
    const int ci = t0 + threadIdx.x;
    const int f = (ci < n_chunks && flag[ch[3 * ci]]) ? 1 : 0;
    const int pa = block_scan_incl(f, sa);
    if (f) chunk[base_c + pa - 1] = ci;
    base_c += sa[AT - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) counts[1] = base_c;
}


// J_cc += sigma[instance] |J_cc| on the diagonal of every unconstrained dof
// (pseudo-transient continuation); Dirichlet components and hanging nodes
// keep their rows
template <typename TJ>
__global__ __launch_bounds__(RT) void k_diag_shift(int n_rows, const int *__restrict__ diag,
                                                   const uint32_t *__restrict__ mask, const int *__restrict__ row_seg,
                                                   const double *__restrict__ sigma, TJ *__restrict__ J)
{
  const int i = blockIdx.x * RT + threadIdx.x;
  if (i >= n_rows * 18) return;
  const int r = i / 18, c = i % 18;
  const uint32_t mk = mask[r];
  if ((mk & HANG_BIT) || ((mk >> c) & 1u)) return;
  TJ *p = J + (size_t)diag[r] * 324 + c * 19;
  const double v = (double)*p;
  *p = (TJ)(v + sigma[row_seg[r]] * fabs(v));
}

// ----------------------------------------- device-side GMRES control (batched)
// One thread per instance.  op 0: start of a restart cycle (beta, g, cycle
// flags); op 1: DGKS flag for column j; op 2: Givens update of column j;
// op 3, the back-substitution for y, has a kernel of its own below.
template <typename T>
__device__ inline void gmres_ctl_op(const GmresCtl<T> &C, int op, int j, int s)
{
  const int m = C.m, ldh = C.ldh;
  double *g = C.g + (size_t)s * (m + 1);
  double *H = C.H + (size_t)s * (m + 1) * m;
  double *cs = C.cs + (size_t)s * m, *sn = C.sn + (size_t)s * m;
  if (op == 0) {
    const double b = C.nrm2[s] > (T)0 ? sqrt((double)C.nrm2[s]) : 0.0;
    C.beta[s] = b;
    const int it = !C.frozen[s] && b > C.target[s];
    C.cyc[s] = it;
    C.conv[s] = !it;
    for (int r = 0; r <= m; ++r) g[r] = 0.0;
    g[0] = b;
    C.jlast[s] = C.jseg[s];
    C.jseg[s] = 0;
    return;
  }
  if (op == 1) {
    C.ro[s] = C.cyc[s] && (double)C.nrm2[s] < C.dgks * (double)C.hs[(size_t)s * ldh + j + 1];
    return;
  }
  if (op == 2) {
    if (!C.cyc[s]) return;
    const double hn = C.nrm2[s] > (T)0 ? sqrt((double)C.nrm2[s]) : 0.0;
    const T *hs = C.hs + (size_t)s * ldh, *hq = C.hq + (size_t)s * ldh;
    const bool ro = C.ro[s] != 0;
    // apply the previous rotations to the new column; the entry carried from
    // rotation r to r+1 stays in a register, the column entries and the
    // rotations are independent loads
    double cur = (double)hs[0] + (ro ? (double)hq[0] : 0.0);
    for (int r = 0; r < j; ++r) {
      const double nxt = (double)hs[r + 1] + (ro ? (double)hq[r + 1] : 0.0);
      const double c_ = cs[r], s_ = sn[r];
      H[(size_t)r * m + j] = c_ * cur + s_ * nxt;
      cur = -s_ * cur + c_ * nxt;
    }
    const double rr = hypot(cur, hn);
    cs[j] = rr > 0 ? cur / rr : 1.0;
    sn[j] = rr > 0 ? hn / rr : 0.0;
    H[(size_t)j * m + j] = rr;
    g[j + 1] = -sn[j] * g[j];
    g[j] = cs[j] * g[j];
    C.jseg[s] = j + 1;
    C.its[s] += 1;
    if (fabs(g[j + 1]) <= C.target[s] || hn == 0.0) C.cyc[s] = 0;
    return;
  }
  // op 3 (the back-substitution) is k_gmres_backsub
}

template <typename T>
__global__ void k_gmres_ctl(GmresCtl<T> C, int op, int j)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < C.S) gmres_ctl_op(C, op, j, s);
}

// Control op 3 (y = R^-1 g on the first jseg columns, zero beyond) with one
// wavefront per instance: each row's sum over the columns already solved is
// formed by the lanes from a coalesced read of the row of R and a wavefront
// sum (one thread per instance, reading its rows serially, took about 0.2 ms
// for a batch of 256 instances at 100 columns).
template <typename T>
__global__ __launch_bounds__(64) void k_gmres_backsub(GmresCtl<T> C)
{
  __shared__ double ys[128];
  const int s = blockIdx.x, lane = threadIdx.x;
  if (s >= C.S) return;
  const int m = C.m, ldh = C.ldh, ji = C.jseg[s];
  const double *g = C.g + (size_t)s * (m + 1);
  const double *H = C.H + (size_t)s * (m + 1) * m;
  T *y = C.y + (size_t)s * ldh;
  for (int r = ji + lane; r < ldh; r += 64) y[r] = (T)0;
  for (int r = ji - 1; r >= 0; --r) {
    double part = 0.0;
    for (int k = r + 1 + lane; k < ji; k += 64) part += H[(size_t)r * m + k] * ys[k];
    part = wave_sum(part);
    if (lane == 0) {
      const double yr = (g[r] - part) / H[(size_t)r * m + r];
      ys[r] = yr;
      y[r] = (T)yr;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");  // ys[r] before the next row reads it
    __syncthreads();
  }
}

// The norm reduction of k_seg_final (one value per instance, in its order)
// and then, on the same workgroup, control op 1 or 2 of that instance: one
// launch instead of two per Gram-Schmidt pass.  The instances visited are the
// listed ones (act) or all; the op of an instance runs whether or not its
// reduction is flagged (op 1 clears the second-pass flag of an instance that
// stopped iterating, op 2 returns at once for it).
template <typename T>
__global__ __launch_bounds__(RT) void k_seg_final_ctl(const int *__restrict__ seg_chunk, int n_seg, ActiveSet act,
                                                      const T *__restrict__ partial, T *norm2,
                                                      const int *__restrict__ flag, GmresCtl<T> C, int op, int j)
{
  __shared__ T sh[RT];
  const int na = act.counts ? act.counts[0] : n_seg;
  for (int a = blockIdx.x; a < na; a += gridDim.x) {
    const int sg = act.counts ? act.seg[a] : a;
    if (!flag || flag[sg]) {
      T acc = 0;
      for (int b = seg_chunk[sg] + threadIdx.x; b < seg_chunk[sg + 1]; b += RT) acc += partial[b];
      const T r = block_reduce(acc, sh);
      if (threadIdx.x == 0) norm2[sg] = r;
    }
    if (threadIdx.x == 0) gmres_ctl_op(C, op, j, sg);  // reads the norm this thread wrote
  }
}

// After the second Gram-Schmidt pass: its norm reduction, control op 2 and
// the normalisation of V_{j+1} in one launch over the chunks (the second
// pass rarely runs; where it did not, the first pass's norm is read).  Every workgroup of an iterating instance forms
// the norm as k_seg_final_ctl does (the chunk partials of the second pass
// where it ran, else the first pass's norm), with the same arithmetic, so all
// of them and the op agree; the workgroup of the instance's first chunk runs
// op 2 (Givens update and stopping test), and every workgroup scales its
// chunk of w.  An instance that the op stops keeps V_{j+1} partly scaled:
// it is not used, as with the separate launches.
template <typename T>
__global__ __launch_bounds__(RT) void k_seg_ctl2_scale(const long long *__restrict__ ch, int n_chunks, ActiveSet act,
                                                       const int *__restrict__ seg_chunk,
                                                       const T *__restrict__ partial, T *norm2, GmresCtl<T> C,
                                                       int j, T *__restrict__ w)
{
  __shared__ T sh[RT];
  __shared__ int s_it, s_ro;
  __shared__ T s_n2;
  __shared__ double s_col[128], s_cs[128], s_sn[128];  // op 2: j < m <= 126
  const int nw = work_count(act, n_chunks);
  for (int wi = blockIdx.x; wi < nw; wi += gridDim.x) {
    const int cb = work_chunk(act, wi);
    const long long sg = ch[3 * cb], beg = ch[3 * cb + 1], end = ch[3 * cb + 2];
    // the flags and the first pass's norm, read once per workgroup with the
    // three loads in flight together: the op of this instance may clear the
    // flag meanwhile, and the branch below holds a block reduction
    if (threadIdx.x == 0) {
      const int it0 = C.cyc[sg], ro0 = C.ro[sg];
      const T n20 = norm2[sg];
      s_it = it0;
      s_ro = ro0;
      s_n2 = n20;
    }
    __syncthreads();
    const bool it = s_it != 0, ro = s_ro != 0;
    T nrm = s_n2;
    __syncthreads();
    if (!it) continue;
    if (ro) {
      T acc = 0;
      for (int b = seg_chunk[sg] + threadIdx.x; b < seg_chunk[sg + 1]; b += RT) acc += partial[b];
      nrm = block_reduce(acc, sh);
    }
    if (cb == seg_chunk[sg]) {
      // op 2 with the new column (both passes summed) and the previous
      // rotations staged in LDS by the workgroup: the rotation loop of one
      // thread then reads LDS, not a chain of global loads (same arithmetic
      // and order as gmres_ctl_op)
      const size_t so = (size_t)sg * C.ldh, sr = (size_t)sg * C.m;
      for (int r = threadIdx.x; r <= j; r += RT)
        s_col[r] = (double)C.hs[so + r] + (ro ? (double)C.hq[so + r] : 0.0);
      for (int r = threadIdx.x; r < j; r += RT) {
        s_cs[r] = C.cs[sr + r];
        s_sn[r] = C.sn[sr + r];
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        if (ro) norm2[sg] = nrm;
        const int m = C.m;
        double *g = C.g + (size_t)sg * (m + 1);
        double *H = C.H + (size_t)sg * (m + 1) * m;
        double *cs = C.cs + sr, *sn = C.sn + sr;
        const double hn = nrm > (T)0 ? sqrt((double)nrm) : 0.0;
        double cur = s_col[0];
        for (int r = 0; r < j; ++r) {
          const double nxt = s_col[r + 1], c_ = s_cs[r], s_ = s_sn[r];
          H[(size_t)r * m + j] = c_ * cur + s_ * nxt;
          cur = -s_ * cur + c_ * nxt;
        }
        const double rr = hypot(cur, hn);
        cs[j] = rr > 0 ? cur / rr : 1.0;
        sn[j] = rr > 0 ? hn / rr : 0.0;
        H[(size_t)j * m + j] = rr;
        g[j + 1] = -sn[j] * g[j];
        g[j] = cs[j] * g[j];
        C.jseg[sg] = j + 1;
        C.its[sg] += 1;
        if (fabs(g[j + 1]) <= C.target[sg] || hn == 0.0) C.cyc[sg] = 0;
      }
    }
    const T a = nrm > (T)0 ? sqrt(nrm) : (T)0;
    const T inv = a != (T)0 ? (T)1 / a : (T)1;
    for (long long idx = beg + threadIdx.x; idx < end; idx += RT) w[idx] = w[idx] * inv;
  }
}


template <typename T>
__global__ __launch_bounds__(RT) void k_fill(size_t n, T v, T *__restrict__ d)
{
  const size_t stride = (size_t)gridDim.x * RT;
  for (size_t idx = (size_t)blockIdx.x * RT + threadIdx.x; idx < n; idx += stride) d[idx] = v;
}

// r = b - w per chunk with partial ||r||^2
template <typename T>
__global__ __launch_bounds__(RT) void k_seg_residual(const long long *__restrict__ ch, const T *__restrict__ b,
                                                     const T *__restrict__ w, T *__restrict__ r,
                                                     T *__restrict__ partial)
{
  __shared__ T sh[RT];
  const long long beg = ch[3 * blockIdx.x + 1], end = ch[3 * blockIdx.x + 2];
  T acc = 0;
  for (long long idx = beg + threadIdx.x; idx < end; idx += RT) {
    const T v = b[idx] - w[idx];
    r[idx] = v;
    acc += v * v;
  }
  const T res = block_reduce(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = res;
}

inline int grid_for(size_t n)
{
  size_t g = (n + RT - 1) / RT;
  if (g > 4096) g = 4096;
  if (g < 1) g = 1;
  return (int)g;
}

}  // namespace

template <typename TJ, typename TX>
hipError_t launch_spmv(int n_rows, const int *row_ptr, const int *col, const TJ *val, const TX *x, TX *y,
                       hipStream_t s, const int *row_seg, const int *seg_flag, ActiveSet act)
{
  if (n_rows <= 0) return hipSuccess;
  const int cap = act.row_cap > 0 ? act.row_cap : ACT_ROW_GRID;
  const int grid = act.counts ? std::max(1, std::min((n_rows + 3) / 4, cap) >> act.shift) : (n_rows + 3) / 4;
  hipLaunchKernelGGL((k_spmv<TJ, TX>), dim3(grid), dim3(256), 0, s, n_rows, row_ptr, col, val, x, y, row_seg,
                     seg_flag, act);
  return hipGetLastError();
}

template <typename TJ>
hipError_t launch_residual_spmv(int n_rows, const int *row_ptr, const int *col, const TJ *val, const double *x,
                                const double *b, double *r, hipStream_t s)
{
  if (n_rows <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_res_spmv<TJ>), dim3((n_rows + 3) / 4), dim3(256), 0, s, n_rows, row_ptr, col, val, x, b, r);
  return hipGetLastError();
}

template <typename TJ>
hipError_t launch_bj_invert(int n_rows, const int *diag_blk, const TJ *val, TJ *Dinv, int *status, hipStream_t s)
{
  if (n_rows <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_bj_invert<TJ>), dim3((n_rows + 11) / 12), dim3(256), 0, s, n_rows, diag_blk, val, Dinv, status);
  return hipGetLastError();
}

template <typename TJ, typename TX>
hipError_t launch_bj_apply(int n_rows, const TJ *Dinv, const TX *x, TX *y, hipStream_t s, const int *row_seg,
                           const int *seg_flag, ActiveSet act)
{
  if (n_rows <= 0) return hipSuccess;
  const int cap = act.row_cap > 0 ? act.row_cap : ACT_ROW_GRID;
  const int g12 = (n_rows + 11) / 12;  // 4 wavefronts of 3 rows
  const int grid = act.counts ? std::max(1, std::min(g12, cap) >> act.shift) : g12;
  hipLaunchKernelGGL((k_bj_apply<TJ, TX>), dim3(grid), dim3(256), 0, s, n_rows, Dinv, x, y, row_seg, seg_flag, act);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_mdot(size_t n, int k, const T *V, size_t ldv, const T *w, T *partial, T *out, bool accumulate,
                       hipStream_t s)
{
  if (k <= 0) return hipSuccess;
  for (int i0 = 0; i0 < k; i0 += KCH)
    hipLaunchKernelGGL((k_mdot_partial<T>), dim3(MDOT_GRID), dim3(RT), 0, s, n, k, i0, V, ldv, w, partial);
  hipLaunchKernelGGL((k_mdot_final<T>), dim3(k), dim3(RT), 0, s, MDOT_GRID, k, partial, out, accumulate ? 1 : 0);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_sub_lincomb(size_t n, int k, const T *V, size_t ldv, const T *h, T *w, hipStream_t s)
{
  hipLaunchKernelGGL((k_sub_lincomb<T>), dim3(grid_for(n)), dim3(RT), 0, s, n, k, V, ldv, h, w);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_sub_lincomb_norm(size_t n, int k, const T *V, size_t ldv, const T *h, T *w, T *partial, T *norm2,
                                   hipStream_t s)
{
  hipLaunchKernelGGL((k_sub_lincomb_norm<T>), dim3(MDOT_GRID), dim3(RT), 0, s, n, k, V, ldv, h, w, partial);
  hipLaunchKernelGGL((k_mdot_final<T>), dim3(1), dim3(RT), 0, s, MDOT_GRID, 1, partial, norm2, 0);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_lincomb(size_t n, int k, const T *V, size_t ldv, const T *y, T *w, hipStream_t s)
{
  hipLaunchKernelGGL((k_lincomb<T>), dim3(grid_for(n)), dim3(RT), 0, s, n, k, V, ldv, y, w);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_scale_by(size_t n, const T *src, const T *scal, T *dst, hipStream_t s)
{
  hipLaunchKernelGGL((k_scale_by<T>), dim3(grid_for(n)), dim3(RT), 0, s, n, src, scal, dst);
  return hipGetLastError();
}

template <typename TX, typename TY>
hipError_t launch_axpy(size_t n, double a, const TX *x, TY *y, hipStream_t s)
{
  hipLaunchKernelGGL((k_axpy<TX, TY>), dim3(grid_for(n)), dim3(RT), 0, s, n, a, x, y);
  return hipGetLastError();
}

template <typename TS, typename TD>
hipError_t launch_convert(size_t n, const TS *src, TD *dst, hipStream_t s)
{
  hipLaunchKernelGGL((k_convert<TS, TD>), dim3(grid_for(n)), dim3(RT), 0, s, n, src, dst);
  return hipGetLastError();
}

template <typename TJ>
hipError_t launch_gather_blocks(int n_rows, const int *blk, const TJ *val, double *dst, hipStream_t s)
{
  const long long n = (long long)n_rows * 324;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_gather_blocks<TJ>), dim3((unsigned)((n + RT - 1) / RT)), dim3(RT), 0, s, n_rows, blk, val, dst);
  return hipGetLastError();
}
template hipError_t launch_gather_blocks<double>(int, const int *, const double *, double *, hipStream_t);
template hipError_t launch_gather_blocks<float>(int, const int *, const float *, double *, hipStream_t);

hipError_t launch_sub(size_t n, const double *b, const double *w, double *dst, hipStream_t s)
{
  hipLaunchKernelGGL(k_sub, dim3(grid_for(n)), dim3(RT), 0, s, n, b, w, dst);
  return hipGetLastError();
}

hipError_t launch_trial(int n_nodes, const double *U, const double *dU, double a, const uint32_t *mask, double *X,
                        hipStream_t s)
{
  const size_t n = (size_t)n_nodes * 18;
  hipLaunchKernelGGL(k_trial, dim3(grid_for(n)), dim3(RT), 0, s, n, U, dU, a, mask, X);
  return hipGetLastError();
}

hipError_t launch_mask_zero(int n_nodes, const uint32_t *mask, double *x, hipStream_t s)
{
  const size_t n = (size_t)n_nodes * 18;
  hipLaunchKernelGGL(k_mask_zero, dim3(grid_for(n)), dim3(RT), 0, s, n, mask, x);
  return hipGetLastError();
}

hipError_t launch_sum(size_t n, const double *v, double *partial, double *out, hipStream_t s)
{
  hipLaunchKernelGGL(k_sum_partial, dim3(MDOT_GRID), dim3(RT), 0, s, n, v, partial);
  hipLaunchKernelGGL((k_mdot_final<double>), dim3(1), dim3(RT), 0, s, MDOT_GRID, 1, partial, out, 0);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_set_scalar(T *p, T v, hipStream_t s)
{
  hipLaunchKernelGGL((k_set_scalar<T>), dim3(1), dim3(64), 0, s, p, v);
  return hipGetLastError();
}

hipError_t launch_read(size_t nbytes, const void *src, float *out, int grid, hipStream_t s)
{
  hipLaunchKernelGGL(k_read16, dim3(grid), dim3(RT), 0, s, nbytes / 16, (const float4 *)src, out);
  return hipGetLastError();
}

hipError_t launch_copy(size_t nbytes, const void *src, void *dst, hipStream_t s)
{
  const size_t n16 = nbytes / 16;
  size_t g = (n16 + RT - 1) / RT;
  if (g > 8192) g = 8192;
  hipLaunchKernelGGL(k_copy16, dim3((int)(g ? g : 1)), dim3(RT), 0, s, n16, (const float4 *)src, (float4 *)dst);
  return hipGetLastError();
}

inline int chunk_grid(const SegInfo &S)
{
  const int cap = S.act.chunk_cap > 0 ? S.act.chunk_cap : ACT_CHUNK_GRID;
  if (S.act.counts) return std::max(1, std::min(S.n_chunks, cap) >> S.act.shift);
  return S.cap > 0 ? std::min(S.n_chunks, S.cap) : S.n_chunks;
}
inline int seg_grid(const SegInfo &S)
{
  return S.act.counts ? std::max(1, std::min(S.n_seg, ACT_SEG_GRID) >> S.act.shift) : S.n_seg;
}

template <typename T>
hipError_t launch_seg_mdot(const SegInfo &S, int k, const T *V, size_t ldv, const T *w, T *partial, T *out, int ldo,
                           int off, bool accumulate, hipStream_t s, const int *flag, bool reduce, const Op1In<T> &O)
{
  if (k <= 0 || S.n_chunks <= 0) return hipSuccess;
  if (O.npart && (!O.norm2 || O.norm2 != O.C.nrm2 || O.npart == partial || reduce)) return hipErrorInvalidValue;
  int gx = S.act.counts ? chunk_grid(S) : (S.n_chunks + 7) / 8 * 8;  // XCD-numbered: a multiple of 8
  if (!S.act.counts && S.cap > 0 && S.cap < S.n_chunks) gx = S.cap;  // grid-stride over the chunks
#define GLH_MDOT(G)                                                                                              \
  hipLaunchKernelGGL((k_seg_mdot_partial<T, G>), dim3(gx, (k + G - 1) / G), dim3(RT), 0, s, S.chunks, S.n_chunks, \
                     S.act, k, V, ldv, w, partial, flag, O)
  if (S.gs == 4) GLH_MDOT(4);
  else if (S.gs == 8) GLH_MDOT(8);
  else if (S.gs == 16) GLH_MDOT(16);
  else return hipErrorInvalidValue;
#undef GLH_MDOT
  if (!reduce) return hipGetLastError();  // the consumer sums the chunk partials itself
  hipLaunchKernelGGL((k_seg_final<T>), dim3(seg_grid(S), k), dim3(RT), 0, s, S.seg_chunk, S.n_seg, S.act, k, partial,
                     out, ldo, off, accumulate ? 1 : 0, flag);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_seg_sub_lincomb_norm(const SegInfo &S, int k, const T *V, size_t ldv, T *h, int ldh, T *w,
                                       T *partial, T *norm2, int ldo, int off, hipStream_t s, const int *flag,
                                       const GmresCtl<T> *ctl, int op, int j, const T *hpart, int hk, bool reduce)
{
  if (S.n_chunks <= 0) return hipSuccess;
  if (hpart && (hk < k || hk > 128 || hpart == partial)) return hipErrorInvalidValue;
#define GLH_SLN(UF)                                                                                              \
  hipLaunchKernelGGL((k_seg_sub_lincomb_norm<T, UF>), dim3(chunk_grid(S)), dim3(RT), 0, s, S.chunks, S.n_chunks, \
                     S.act, k, V, ldv, h, ldh, w, partial, flag, S.seg_chunk, hpart, hk)
  if (S.gs == 4) GLH_SLN(4);
  else if (S.gs == 8) GLH_SLN(8);
  else if (S.gs == 16) GLH_SLN(16);
  else return hipErrorInvalidValue;
#undef GLH_SLN
  if (!reduce) return hipGetLastError();  // the norm partials go to a later launch
  if (ctl) {
    if (ldo != 1 || off != 0 || norm2 != ctl->nrm2) return hipErrorInvalidValue;  // the op reads the norms it gets
    hipLaunchKernelGGL((k_seg_final_ctl<T>), dim3(seg_grid(S)), dim3(RT), 0, s, S.seg_chunk, S.n_seg, S.act, partial,
                       norm2, flag, *ctl, op, j);
  } else {
    hipLaunchKernelGGL((k_seg_final<T>), dim3(seg_grid(S), 1), dim3(RT), 0, s, S.seg_chunk, S.n_seg, S.act, 1,
                       partial, norm2, ldo, off, 0, flag);
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_seg_ctl2_scale(const SegInfo &S, const T *partial, T *norm2, const GmresCtl<T> &C, int j, T *w,
                                 hipStream_t s)
{
  if (S.n_chunks <= 0) return hipSuccess;
  if (norm2 != C.nrm2) return hipErrorInvalidValue;  // the op reads the norms it gets
  hipLaunchKernelGGL((k_seg_ctl2_scale<T>), dim3(chunk_grid(S)), dim3(RT), 0, s, S.chunks, S.n_chunks, S.act,
                     S.seg_chunk, partial, norm2, C, j, w);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_seg_lincomb(const SegInfo &S, int k, const T *V, size_t ldv, const T *y, int ldy, T *w,
                              hipStream_t s, const int *kseg)
{
  if (S.n_chunks <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_seg_lincomb<T>), dim3(chunk_grid(S)), dim3(RT), 0, s, S.chunks, S.n_chunks, S.act, k, V, ldv,
                     y, ldy, w, kseg);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_seg_scale(const SegInfo &S, const T *src, const T *scal, int lds, T *dst, hipStream_t s,
                            const int *flag, bool sqrt_of_scale)
{
  if (S.n_chunks <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_seg_scale<T>), dim3(chunk_grid(S)), dim3(RT), 0, s, S.chunks, S.n_chunks, S.act, src, scal,
                     lds, dst, flag, sqrt_of_scale ? 1 : 0);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_seg_add(const SegInfo &S, const T *x, T *y, hipStream_t s)
{
  if (S.n_chunks <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_seg_add<T>), dim3(chunk_grid(S)), dim3(RT), 0, s, S.chunks, S.n_chunks, S.act, x, y);
  return hipGetLastError();
}

hipError_t launch_build_active(const SegInfo &S, const int *flag, const int *node_off, int *counts, int *seg,
                               int *rpref, int *chunk, hipStream_t s)
{
  hipLaunchKernelGGL(k_build_active, dim3(1), dim3(AT), 0, s, S.n_seg, S.n_chunks, S.chunks, flag, node_off, counts,
                     seg, rpref, chunk);
  return hipGetLastError();
}

hipError_t launch_seg_trial(const SegInfo &S, const double *U, const double *dU, const double *alpha,
                            const uint32_t *mask, double *X, hipStream_t s)
{
  if (S.n_chunks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_seg_trial, dim3(S.n_chunks), dim3(RT), 0, s, S.chunks, U, dU, alpha, mask, X);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_seg_zero(const SegInfo &S, const int *flag, T *v, hipStream_t s)
{
  if (S.n_chunks <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_seg_zero<T>), dim3(S.n_chunks), dim3(RT), 0, s, S.chunks, flag, v);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_gmres_ctl(const GmresCtl<T> &C, int op, int j, hipStream_t s)
{
  if (op == 3) hipLaunchKernelGGL((k_gmres_backsub<T>), dim3(C.S), dim3(64), 0, s, C);
  else hipLaunchKernelGGL((k_gmres_ctl<T>), dim3((C.S + 63) / 64), dim3(64), 0, s, C, op, j);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_fill(size_t n, T v, T *d, hipStream_t s)
{
  hipLaunchKernelGGL((k_fill<T>), dim3(grid_for(n)), dim3(RT), 0, s, n, v, d);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_seg_residual(const SegInfo &S, const T *b, const T *w, T *r, T *partial, T *norm2, hipStream_t s)
{
  hipLaunchKernelGGL((k_seg_residual<T>), dim3(S.n_chunks), dim3(RT), 0, s, S.chunks, b, w, r, partial);
  hipLaunchKernelGGL((k_seg_final<T>), dim3(S.n_seg, 1), dim3(RT), 0, s, S.seg_chunk, S.n_seg, ActiveSet{}, 1, partial,
                     norm2, 1, 0, 0, (const int *)nullptr);
  return hipGetLastError();
}

#define GLH_SEG_INST(T)                                                                                         \
  template hipError_t launch_seg_mdot<T>(const SegInfo &, int, const T *, size_t, const T *, T *, T *, int, int,   \
                                         bool, hipStream_t, const int *, bool, const Op1In<T> &);                  \
  template hipError_t launch_seg_sub_lincomb_norm<T>(const SegInfo &, int, const T *, size_t, T *, int, T *,       \
                                                     T *, T *, int, int, hipStream_t, const int *,                 \
                                                     const GmresCtl<T> *, int, int, const T *, int, bool);         \
  template hipError_t launch_seg_ctl2_scale<T>(const SegInfo &, const T *, T *, const GmresCtl<T> &, int, T *,     \
                                               hipStream_t);                                                       \
  template hipError_t launch_seg_lincomb<T>(const SegInfo &, int, const T *, size_t, const T *, int, T *,          \
                                            hipStream_t, const int *);                                             \
  template hipError_t launch_seg_scale<T>(const SegInfo &, const T *, const T *, int, T *, hipStream_t,            \
                                          const int *, bool);                                                      \
  template hipError_t launch_seg_zero<T>(const SegInfo &, const int *, T *, hipStream_t);                          \
  template hipError_t launch_gmres_ctl<T>(const GmresCtl<T> &, int, int, hipStream_t);                             \
  template hipError_t launch_fill<T>(size_t, T, T *, hipStream_t);                                                 \
  template hipError_t launch_seg_residual<T>(const SegInfo &, const T *, const T *, T *, T *, T *, hipStream_t);  \
  template hipError_t launch_seg_add<T>(const SegInfo &, const T *, T *, hipStream_t);
GLH_SEG_INST(double)
GLH_SEG_INST(float)
#undef GLH_SEG_INST
template hipError_t launch_fill<int>(size_t, int, int *, hipStream_t);

template <typename TJ>
hipError_t launch_diag_shift(int n_rows, const int *diag, const uint32_t *mask, const int *row_seg,
                             const double *sigma, TJ *J, hipStream_t s)
{
  if (n_rows <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_diag_shift<TJ>), dim3((n_rows * 18 + RT - 1) / RT), dim3(RT), 0, s, n_rows, diag, mask, row_seg,
                     sigma, J);
  return hipGetLastError();
}
template hipError_t launch_diag_shift<double>(int, const int *, const uint32_t *, const int *, const double *, double *,
                                              hipStream_t);
template hipError_t launch_diag_shift<float>(int, const int *, const uint32_t *, const int *, const double *, float *,
                                             hipStream_t);

// explicit instantiations
template hipError_t launch_set_scalar<double>(double *, double, hipStream_t);
template hipError_t launch_set_scalar<float>(float *, float, hipStream_t);
template hipError_t launch_axpy<float, float>(size_t, double, const float *, float *, hipStream_t);
template hipError_t launch_spmv<double, double>(int, const int *, const int *, const double *, const double *,
                                                double *, hipStream_t, const int *, const int *, ActiveSet);
template hipError_t launch_spmv<float, float>(int, const int *, const int *, const float *, const float *, float *,
                                              hipStream_t, const int *, const int *, ActiveSet);
template hipError_t launch_residual_spmv<double>(int, const int *, const int *, const double *, const double *,
                                                 const double *, double *, hipStream_t);
template hipError_t launch_residual_spmv<float>(int, const int *, const int *, const float *, const double *,
                                                const double *, double *, hipStream_t);
template hipError_t launch_bj_invert<double>(int, const int *, const double *, double *, int *, hipStream_t);
template hipError_t launch_bj_invert<float>(int, const int *, const float *, float *, int *, hipStream_t);
template hipError_t launch_bj_apply<double, double>(int, const double *, const double *, double *, hipStream_t,
                                                    const int *, const int *, ActiveSet);
template hipError_t launch_bj_apply<float, float>(int, const float *, const float *, float *, hipStream_t,
                                                  const int *, const int *, ActiveSet);
template hipError_t launch_mdot<double>(size_t, int, const double *, size_t, const double *, double *, double *, bool,
                                        hipStream_t);
template hipError_t launch_mdot<float>(size_t, int, const float *, size_t, const float *, float *, float *, bool,
                                       hipStream_t);
template hipError_t launch_sub_lincomb<double>(size_t, int, const double *, size_t, const double *, double *,
                                               hipStream_t);
template hipError_t launch_sub_lincomb<float>(size_t, int, const float *, size_t, const float *, float *, hipStream_t);
template hipError_t launch_sub_lincomb_norm<double>(size_t, int, const double *, size_t, const double *, double *,
                                                    double *, double *, hipStream_t);
template hipError_t launch_sub_lincomb_norm<float>(size_t, int, const float *, size_t, const float *, float *, float *,
                                                   float *, hipStream_t);
template hipError_t launch_lincomb<double>(size_t, int, const double *, size_t, const double *, double *, hipStream_t);
template hipError_t launch_lincomb<float>(size_t, int, const float *, size_t, const float *, float *, hipStream_t);
template hipError_t launch_scale_by<double>(size_t, const double *, const double *, double *, hipStream_t);
template hipError_t launch_scale_by<float>(size_t, const float *, const float *, float *, hipStream_t);
template hipError_t launch_axpy<double, double>(size_t, double, const double *, double *, hipStream_t);
template hipError_t launch_axpy<float, double>(size_t, double, const float *, double *, hipStream_t);
template hipError_t launch_convert<double, float>(size_t, const double *, float *, hipStream_t);
template hipError_t launch_convert<float, double>(size_t, const float *, double *, hipStream_t);

}  // namespace glh
