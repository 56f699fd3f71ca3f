// Coordinate-MLP warm start on gfx950: hand-written BF16 kernels with FP32
// accumulation for training (forward, backward, Adam) and inference.
//
// Network: x (6 inputs) -> Fourier features [sin, cos](2 pi x B), B a fixed
// 6 x F matrix -> `depth` hidden layers of width W with SiLU -> n_out linear
// outputs.  Activations and weights enter the matrix cores as BF16; every
// product accumulates in FP32 (v_mfma_f32_16x16x32_bf16); master weights,
// gradients and Adam moments are FP32.
//
// GEMM: C (FP32) = A B with A (M x K) and B (K x N) in BF16, 64 x 64 tiles,
// 4 wavefronts of 2 x 2 MFMA 16 x 16 blocks, M, N multiples of 64 and K a
// multiple of 32 (the host pads).  The network's products run in two
// kernels: k_gemm_k128 (both operands contiguous along k: the forward,
// output-layer and input-gradient products with their epilogues) and
// k_gemm_wgrad (both contiguous along their other index: the weight
// gradients, split along the batch); k_gemm serves the other layouts of
// the plain product, 32-wide k steps with transposing LDS stores.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mlp_dev.hpp"

namespace glh {
namespace mlp {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int BM = 64, BN = 64, BK = 32, LDK = BK + 8;  // LDS rows padded by 16 B

__device__ inline bf16 to_bf16(float f)
{
  // round to nearest even on the upper 16 bits (finite inputs)
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return bf16{(uint16_t)(u >> 16)};
}
__device__ inline float from_bf16(bf16 h) { return __uint_as_float((uint32_t)h.bits << 16); }

// A(m,k) = A[m*sam + k*sak] with A_K (sak == 1) or A_M (sam == 1);
// B(k,n) = B[k*sbk + n*sbn] with B_K (sbk == 1) or B_N (sbn == 1).
// Sum over the k range [blockIdx.z * kchunk, +kchunk) of A(m,k) B(k,n), then
// C[m*ldc + n] (+)= it at C + blockIdx.z * c_split (GEMM_STORE), with the
// column sums of B for CS.
template <bool A_K, bool B_K, bool CS>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int kchunk, const bf16 *__restrict__ A, long long sa,
                                              const bf16 *__restrict__ B, long long sb, float *__restrict__ C,
                                              int ldc, long long c_split, int accumulate, GemmEpi E)
{
  __shared__ __attribute__((aligned(16))) uint16_t As[BM][LDK];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[BN][LDK];
  __shared__ float red[CS ? 32 : 1][BN];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const int kbeg = blockIdx.z * kchunk;
  const uint16_t *Ag = reinterpret_cast<const uint16_t *>(A);
  const uint16_t *Bg = reinterpret_cast<const uint16_t *>(B);
  // column sums of B (row block 0 only): thread t sums the 8 columns it stages
  const bool cs = CS && blockIdx.y == 0;
  float csum[CS ? 8 : 1];
#pragma unroll
  for (int i = 0; i < (CS ? 8 : 1); ++i) csum[i] = 0.f;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  for (int k0 = kbeg; k0 < kbeg + kchunk; k0 += BK) {
    // stage A (64 x 32) and B (32 x 64) tiles, 8 values per thread each
    if (A_K) {
      const int m = t >> 2, kk = (t & 3) * 8;
      const uint4 v = *reinterpret_cast<const uint4 *>(Ag + (size_t)(m0 + m) * sa + (k0 + kk));
      *reinterpret_cast<uint4 *>(&As[m][kk]) = v;
    } else {
      const int kk = t >> 3, m = (t & 7) * 8;
      const uint4 v = *reinterpret_cast<const uint4 *>(Ag + (size_t)(k0 + kk) * sa + (m0 + m));
      const uint16_t *e = reinterpret_cast<const uint16_t *>(&v);
#pragma unroll
      for (int i = 0; i < 8; ++i) As[m + i][kk] = e[i];
    }
    if (B_K) {
      const int n = t >> 2, kk = (t & 3) * 8;
      const uint4 v = *reinterpret_cast<const uint4 *>(Bg + (size_t)(n0 + n) * sb + (k0 + kk));
      *reinterpret_cast<uint4 *>(&Bs[n][kk]) = v;
    } else {
      const int kk = t >> 3, n = (t & 7) * 8;
      const uint4 v = *reinterpret_cast<const uint4 *>(Bg + (size_t)(k0 + kk) * sb + (n0 + n));
      const uint16_t *e = reinterpret_cast<const uint16_t *>(&v);
#pragma unroll
      for (int i = 0; i < 8; ++i) Bs[n + i][kk] = e[i];
      if (CS && cs) {
#pragma unroll
        for (int i = 0; i < 8; ++i) csum[i] += __uint_as_float((uint32_t)e[i] << 16);
      }
    }
    __syncthreads();
    // lane l: row/col l % 16 of the 16 x 16 block, k = 8 (l / 16) .. + 7
    const int r = lane & 15, kq = (lane >> 4) * 8;
    bf16x8 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const uint4 va = *reinterpret_cast<const uint4 *>(&As[wm + i * 16 + r][kq]);
      a[i] = __builtin_bit_cast(bf16x8, va);
      const uint4 vb = *reinterpret_cast<const uint4 *>(&Bs[wn + i * 16 + r][kq]);
      b[i] = __builtin_bit_cast(bf16x8, vb);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  if (CS && cs) {  // the 32 threads staging one group of 8 columns, summed in row order
    const int kk = t >> 3, n = (t & 7) * 8;
#pragma unroll
    for (int i = 0; i < 8; ++i) red[kk][n + i] = csum[i];
    __syncthreads();
    if (t < BN) {
      float s = 0.f;
      for (int q = 0; q < 32; ++q) s += red[q][t];
      E.colsum[(size_t)blockIdx.z * N + n0 + t] = s;
    }
  }
  // D layout: lane l holds rows 4 (l / 16) + q, column l % 16
  float *Cz = C + (size_t)blockIdx.z * c_split;
  const int cc = lane & 15, rq = (lane >> 4) * 4;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int m = m0 + wm + i * 16 + rq + q, n = n0 + wn + j * 16 + cc;
        const float v = acc[i][j][q];
        float *p = Cz + (size_t)m * ldc + n;
        *p = accumulate ? *p + v : v;
      }
}

// The weight gradient C = A B with A(m, k) = A[k*sa + m] and B(k, n) =
// B[k*sb + n] (both operands contiguous along their non-reduced index: the
// activations and dZ, k the batch), GEMM_STORE per split: 64 x 64 tiles as
// k_gemm, staged 128 values of k at a time.  Thread t < 128 stages the 8 x 8
// block (k block t / 8, m block t % 8) of A and thread 128 + t the same of
// B: eight 16-byte loads of 8 values each along m, transposed in registers,
// eight 16-byte LDS stores of 8 values each along k (where k_gemm's
// transposing path stores single 16-bit values).  With CS, row block 0 also
// writes the column sums of B over its k range to colsum[z*N + n].
constexpr int WK = 128, WKP = WK + 8;
template <bool CS>
__global__ __launch_bounds__(256) void k_gemm_wgrad(int N, int kchunk, const bf16 *__restrict__ A, long long sa,
                                                    const bf16 *__restrict__ B, long long sb, float *__restrict__ C,
                                                    int ldc, long long c_split, float *__restrict__ colsum)
{
  __shared__ __attribute__((aligned(16))) uint16_t As[BM][WKP];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[BN][WKP];
  __shared__ float red[CS ? WK / 8 : 1][BN];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const int kbeg = blockIdx.z * kchunk;
  const bool isB = t >= 128;
  const int tb = t & 127, kb = tb >> 3, cb = (tb & 7) * 8;  // k block, column (m or n) block
  const uint16_t *src = reinterpret_cast<const uint16_t *>(isB ? B : A);
  const long long ss = isB ? sb : sa;
  const int c0 = isB ? n0 : m0;
  uint16_t(*dst)[WKP] = isB ? Bs : As;
  const bool cs = CS && isB && blockIdx.y == 0;
  float csum[CS ? 8 : 1];
#pragma unroll
  for (int i = 0; i < (CS ? 8 : 1); ++i) csum[i] = 0.f;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  const int r = lane & 15, kq = (lane >> 4) * 8;
  for (int k0 = kbeg; k0 < kbeg + kchunk; k0 += WK) {
    const int kw = kbeg + kchunk - k0 < WK ? kbeg + kchunk - k0 : WK;
    if (kb * 8 < kw) {
      uint4 x[8];
#pragma unroll
      for (int e = 0; e < 8; ++e)
        x[e] = *reinterpret_cast<const uint4 *>(src + (size_t)(k0 + kb * 8 + e) * ss + c0 + cb);
      // 8 x 8 transpose: column i of the block (8 values along k) from the
      // 16-bit halves of word i / 2 of the eight rows
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        uint32_t wd[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const uint32_t v = (i >> 1) == 0 ? x[e].x : (i >> 1) == 1 ? x[e].y : (i >> 1) == 2 ? x[e].z : x[e].w;
          wd[e] = (i & 1) ? v >> 16 : v & 0xffffu;
        }
        const uint4 col = {wd[0] | (wd[1] << 16), wd[2] | (wd[3] << 16), wd[4] | (wd[5] << 16), wd[6] | (wd[7] << 16)};
        *reinterpret_cast<uint4 *>(&dst[cb + i][kb * 8]) = col;
        if (CS && cs) {
#pragma unroll
          for (int e = 0; e < 8; ++e) csum[i] += __uint_as_float(wd[e] << 16);
        }
      }
    }
    __syncthreads();
    for (int ks = 0; ks < kw; ks += BK) {
      bf16x8 a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        a[i] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(&As[wm + i * 16 + r][ks + kq]));
        b[i] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(&Bs[wn + i * 16 + r][ks + kq]));
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  if (CS && blockIdx.y == 0) {  // the 16 k blocks of each column, summed in k order
    if (cs) {
#pragma unroll
      for (int i = 0; i < 8; ++i) red[kb][cb + i] = csum[i];
    }
    __syncthreads();
    if (t < BN) {
      float s = 0.f;
      for (int q = 0; q < WK / 8; ++q) s += red[q][t];
      colsum[(size_t)blockIdx.z * N + n0 + t] = s;
    }
  }
  float *Cz = C + (size_t)blockIdx.z * c_split;
  const int cc = lane & 15, rq = (lane >> 4) * 4;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q)
        Cz[(size_t)(m0 + wm + i * 16 + rq + q) * ldc + n0 + wn + j * 16 + cc] = acc[i][j][q];
}

// The forward, output-layer and input-gradient products, C = A B with both
// operands contiguous along k (A(m, k) = A[m*sa + k], B(k, n) = B[n*sb + k])
// and the epilogues GEMM_BIAS_SILU, GEMM_BIAS and GEMM_SILU_BACK (GemmEpi):
// 64 x 64 tiles staged 128 values of k at a time, two barriers per 128
// values.  A full stage issues its eight 16-byte loads per thread without
// conditions, so they are all in flight at once (with a per-load condition
// the stage took twice as long); a last, partial stage loads what it holds.
// The epilogue goes through a 64 x 64 FP32 tile in the staging LDS, so that
// Z, C and H are read and written by rows, 4 columns per thread.
template <int EPI>
__global__ __launch_bounds__(256) void k_gemm_k128(int K, const bf16 *__restrict__ A, long long sa,
                                                   const bf16 *__restrict__ B, long long sb, float *__restrict__ C,
                                                   int ldc, GemmEpi E)
{
  // the two stages of A and B, then the 64 x 64 FP32 result tile
  constexpr int TPC = BN + 4;
  static_assert(BM * TPC * 4 <= 2 * BM * WKP * 2, "result tile within the staging buffers");
  __shared__ __attribute__((aligned(16))) uint16_t SM[2 * BM * WKP];
  uint16_t(*As)[WKP] = reinterpret_cast<uint16_t(*)[WKP]>(SM);
  uint16_t(*Bs)[WKP] = reinterpret_cast<uint16_t(*)[WKP]>(SM + BM * WKP);
  float(*Tt)[TPC] = reinterpret_cast<float(*)[TPC]>(SM);
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const uint16_t *Ag = reinterpret_cast<const uint16_t *>(A);
  const uint16_t *Bg = reinterpret_cast<const uint16_t *>(B);
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  const int r = lane & 15, kq = (lane >> 4) * 8;
  for (int k0 = 0; k0 < K; k0 += WK) {
    const int kw = K - k0 < WK ? K - k0 : WK;
    uint4 va[4], vb[4];
    if (kw == WK) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int p = t + 256 * u, row = p >> 4, c = (p & 15) * 8;
        va[u] = *reinterpret_cast<const uint4 *>(Ag + (size_t)(m0 + row) * sa + k0 + c);
        vb[u] = *reinterpret_cast<const uint4 *>(Bg + (size_t)(n0 + row) * sb + k0 + c);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int p = t + 256 * u, row = p >> 4, c = (p & 15) * 8;
        *reinterpret_cast<uint4 *>(&As[row][c]) = va[u];
        *reinterpret_cast<uint4 *>(&Bs[row][c]) = vb[u];
      }
    } else {
      for (int u = 0; u < 4; ++u) {
        const int p = t + 256 * u, row = p >> 4, c = (p & 15) * 8;
        if (c < kw) {
          *reinterpret_cast<uint4 *>(&As[row][c]) = *reinterpret_cast<const uint4 *>(Ag + (size_t)(m0 + row) * sa + k0 + c);
          *reinterpret_cast<uint4 *>(&Bs[row][c]) = *reinterpret_cast<const uint4 *>(Bg + (size_t)(n0 + row) * sb + k0 + c);
        }
      }
    }
    __syncthreads();
    for (int ks = 0; ks < kw; ks += BK) {
      bf16x8 a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        a[i] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(&As[wm + i * 16 + r][ks + kq]));
        b[i] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(&Bs[wn + i * 16 + r][ks + kq]));
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  // epilogue by rows: the accumulators (lane l: rows 4 (l / 16) + q of a
  // block, column l % 16) go to the tile, then each thread takes 4
  // consecutive columns of a row: 16-byte FP32 and 8-byte BF16 accesses
  const int cc = lane & 15, rq = (lane >> 4) * 4;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) Tt[wm + i * 16 + rq + q][wn + j * 16 + cc] = acc[i][j][q];
  __syncthreads();
  const int c4 = (t & 15) * 4, n = n0 + c4;
#pragma unroll
  for (int pass = 0; pass < 4; ++pass) {
    const int row = (t >> 4) + 16 * pass, m = m0 + row;
    const float4 v4 = *reinterpret_cast<const float4 *>(&Tt[row][c4]);
    float o[4] = {v4.x, v4.y, v4.z, v4.w};
    if (EPI == GEMM_BIAS) {
#pragma unroll
      for (int c = 0; c < 4; ++c)
        if (n + c < E.nbias) o[c] += E.bias[n + c];
      *reinterpret_cast<float4 *>(C + (size_t)m * ldc + n) = float4{o[0], o[1], o[2], o[3]};
    } else {
      if (EPI == GEMM_BIAS_SILU) {
        float z[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) z[c] = o[c] + E.bias[n + c];
        *reinterpret_cast<float4 *>(E.Z + (size_t)m * E.ldz + n) = float4{z[0], z[1], z[2], z[3]};
#pragma unroll
        for (int c = 0; c < 4; ++c) o[c] = z[c] / (1.f + expf(-z[c]));
      } else {
        const float4 z4 = *reinterpret_cast<const float4 *>(E.Z + (size_t)m * E.ldz + n);
        const float z[4] = {z4.x, z4.y, z4.z, z4.w};
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const float sg = 1.f / (1.f + expf(-z[c]));
          o[c] = o[c] * sg * (1.f + z[c] * (1.f - sg));
        }
      }
      const uint2 pk = {(uint32_t)to_bf16(o[0]).bits | ((uint32_t)to_bf16(o[1]).bits << 16),
                        (uint32_t)to_bf16(o[2]).bits | ((uint32_t)to_bf16(o[3]).bits << 16)};
      *reinterpret_cast<uint2 *>(reinterpret_cast<uint16_t *>(E.H) + (size_t)m * E.ldh + n) = pk;
    }
  }
}

constexpr int ET = 256;

__global__ __launch_bounds__(ET) void k_fourier(int n, int n_real, const float *__restrict__ X,
                                                const float *__restrict__ Bf, int F, bf16 *__restrict__ H)
{
  const long long i = (long long)blockIdx.x * ET + threadIdx.x;
  if (i >= (long long)n * F) return;
  const int s = (int)(i / F), j = (int)(i % F);
  float a = 0.f;
  if (s < n_real)
    for (int d = 0; d < 6; ++d) a += X[(size_t)s * 6 + d] * Bf[d * F + j];
  const float ang = 6.283185307179586f * a;
  H[(size_t)s * 2 * F + j] = to_bf16(s < n_real ? sinf(ang) : 0.f);
  H[(size_t)s * 2 * F + F + j] = to_bf16(s < n_real ? cosf(ang) : 0.f);
}

// mean-squared error over the n_real x n_out real outputs: dY = bf16(2 (Y - T)
// / (n_real n_out)) (0 in the padding), per-block partial sums of (Y - T)^2
__global__ __launch_bounds__(ET) void k_mse(int n, int n_real, int n_out, int ldy, const float *__restrict__ Y,
                                            const float *__restrict__ T, bf16 *__restrict__ dY,
                                            float *__restrict__ partial)
{
  __shared__ float sh[ET];
  const long long i = (long long)blockIdx.x * ET + threadIdx.x;
  float e2 = 0.f;
  if (i < (long long)n * ldy) {
    const int s = (int)(i / ldy), o = (int)(i % ldy);
    float g = 0.f;
    if (s < n_real && o < n_out) {
      const float d = Y[i] - T[(size_t)s * n_out + o];
      e2 = d * d;
      g = 2.f * d / ((float)n_real * (float)n_out);
    }
    dY[i] = to_bf16(g);
  }
  sh[threadIdx.x] = e2;
  __syncthreads();
  for (int st = ET / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// out[0] = sum of n values, one workgroup, fixed order (strided sums, then a tree)
__global__ __launch_bounds__(ET) void k_reduce1(int n, const float *__restrict__ in, float *__restrict__ out)
{
  __shared__ float sh[ET];
  float s = 0.f;
  int i = threadIdx.x;
  for (; i + 7 * ET < n; i += 8 * ET) {  // 8 loads in flight, summed in the same order
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = in[i + u * ET];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += v[u];
  }
  for (; i < n; i += ET) s += in[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int st = ET / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

struct WbPtrs {
  bf16 *p[MLP_MAX_LAYERS];
};

// layer of flat parameter i (layers in order, each its weights then its biases)
__device__ inline int param_layer(const GradLayout &L, long long i)
{
  int l = 0;
  while (l + 1 < L.n_layers && i >= L.woff[l + 1]) ++l;
  return l;
}

// G[i] = sum over the splits z (in order) of the partials of parameter i: the
// weight-gradient GEMM's part[poff + z K Np + k Np + n], or for a bias the
// column sums cpart[coff + z Np + n]
__global__ __launch_bounds__(ET) void k_grad_reduce(long long n, GradLayout L, const float *__restrict__ part,
                                                    const float *__restrict__ cpart, float *__restrict__ G)
{
  const long long i = (long long)blockIdx.x * ET + threadIdx.x;
  if (i >= n) return;
  const int l = param_layer(L, i);
  const float *src;
  long long stride;
  if (i < L.boff[l]) {
    const long long r = i - L.woff[l];
    const int k = (int)(r / L.N[l]), c = (int)(r % L.N[l]);
    src = part + L.poff[l] + (long long)k * L.Np[l] + c;
    stride = (long long)L.K[l] * L.Np[l];
  } else {
    src = cpart + L.coff[l] + (i - L.boff[l]);
    stride = L.Np[l];
  }
  // the splits in order; 8 loads in flight at a time
  float s = 0.f;
  int z = 0;
  for (; z + 8 <= L.splits; z += 8) {
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = src[(size_t)(z + u) * stride];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += v[u];
  }
  for (; z < L.splits; ++z) s += src[(size_t)z * stride];
  G[i] = s;
}

// Adam on the FP32 master parameters (bias-corrected, step t >= 1); a weight
// also refreshes its BF16 copies Wb[l][k Np + n] and WbT[l][n K + k] (the
// padding stays 0)
__global__ __launch_bounds__(ET) void k_adam_pack(long long n, float *__restrict__ p, const float *__restrict__ g,
                                                  float *__restrict__ m, float *__restrict__ v, float lr, float b1,
                                                  float b2, float eps, const int *__restrict__ step, GradLayout L,
                                                  WbPtrs Wb, WbPtrs WbT)
{
  const long long i = (long long)blockIdx.x * ET + threadIdx.x;
  if (i >= n) return;
  const int t = step[0];
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi, vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / (1.f - powf(b1, (float)t)), vh = vi / (1.f - powf(b2, (float)t));
  const float pn = p[i] - lr * mh / (sqrtf(vh) + eps);
  p[i] = pn;
  const int l = param_layer(L, i);
  if (i < L.boff[l]) {
    const long long r = i - L.woff[l], k = r / L.N[l], c = r % L.N[l];
    const bf16 h = to_bf16(pn);
    Wb.p[l][k * L.Np[l] + c] = h;
    WbT.p[l][c * L.K[l] + k] = h;
  }
}

// BF16 copies of a K x N FP32 weight: K x Np and its transpose Np x K
// (zeros beyond N)
__global__ __launch_bounds__(ET) void k_pack_w(int K, int N, int Np, const float *__restrict__ W,
                                               bf16 *__restrict__ Wb, bf16 *__restrict__ WbT)
{
  const long long i = (long long)blockIdx.x * ET + threadIdx.x;
  if (i >= (long long)K * Np) return;
  const int k = (int)(i / Np), c = (int)(i % Np);
  const bf16 h = to_bf16(c < N ? W[(size_t)k * N + c] : 0.f);
  Wb[i] = h;
  WbT[(size_t)c * K + k] = h;
}

// random training batch: sample i of step t is point hash(seed, t, i) mod n_data
__device__ inline uint64_t mix64(uint64_t z)
{
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
__global__ __launch_bounds__(ET) void k_gather(int nb, long long n_data, uint64_t seed, const int *__restrict__ step,
                                               const float *__restrict__ X, const float *__restrict__ Y, int n_out,
                                               float *__restrict__ Xb, float *__restrict__ Yb)
{
  const int i = blockIdx.x * ET + threadIdx.x;
  if (i >= nb) return;
  const uint64_t h = mix64(seed ^ mix64(((uint64_t)(uint32_t)step[0] << 32) | (uint32_t)i));
  const long long j = (long long)(h % (uint64_t)n_data);
  for (int d = 0; d < 6; ++d) Xb[(size_t)i * 6 + d] = X[(size_t)j * 6 + d];
  for (int o = 0; o < n_out; ++o) Yb[(size_t)i * n_out + o] = Y[(size_t)j * n_out + o];
}

// Y[s, o] = Yp[s*ldy + o] for s < n, o < n_out
__global__ __launch_bounds__(ET) void k_take_out(int n, int n_out, int ldy, const float *__restrict__ Yp,
                                                 float *__restrict__ Y)
{
  const long long i = (long long)blockIdx.x * ET + threadIdx.x;
  if (i >= (long long)n * n_out) return;
  Y[i] = Yp[(size_t)(i / n_out) * ldy + (i % n_out)];
}

__global__ void k_step_inc(int *step)
{
  if (threadIdx.x == 0) step[0] += 1;
}

inline unsigned grid1(long long n) { return (unsigned)((n + ET - 1) / ET); }

}  // namespace

hipError_t gemm(int M, int N, int K, const bf16 *A, bool a_kcontig, long long sa, const bf16 *B, bool b_kcontig,
                long long sb, float *C, int ldc, bool accumulate, int splits, long long c_split, hipStream_t s,
                const GemmEpi &E)
{
  if (M % BM || N % BN || K % (BK * splits)) return hipErrorInvalidValue;
  const dim3 grid(N / BN, M / BM, splits);
  const int kchunk = K / splits;
  const int acc = accumulate ? 1 : 0;
#define GLH_GEMM(AK, BK_, CS) \
  hipLaunchKernelGGL((k_gemm<AK, BK_, CS>), grid, dim3(256), 0, s, M, N, kchunk, A, sa, B, sb, C, ldc, c_split, acc, E)
  // the combinations the network uses
#define GLH_GEMM128(EPI) \
  hipLaunchKernelGGL((k_gemm_k128<EPI>), grid, dim3(256), 0, s, K, A, sa, B, sb, C, ldc, E)
  // the epilogue products take both operands contiguous along k, one split
  if (E.mode != GEMM_STORE && (!a_kcontig || !b_kcontig || splits != 1 || accumulate)) return hipErrorInvalidValue;
  // the epilogue's 16-byte FP32 and 8-byte BF16 row accesses
  if ((E.mode == GEMM_BIAS && ldc % 4) || ((E.mode == GEMM_BIAS_SILU || E.mode == GEMM_SILU_BACK) && (E.ldz % 4 || E.ldh % 4)))
    return hipErrorInvalidValue;
  if (E.mode == GEMM_BIAS_SILU) GLH_GEMM128(GEMM_BIAS_SILU);
  else if (E.mode == GEMM_BIAS) GLH_GEMM128(GEMM_BIAS);
  else if (E.mode == GEMM_SILU_BACK) GLH_GEMM128(GEMM_SILU_BACK);
#undef GLH_GEMM128
  else if (E.mode != GEMM_STORE) return hipErrorInvalidValue;
  else if (!a_kcontig && !b_kcontig && !accumulate) {  // the weight gradients
    if (E.colsum)
      hipLaunchKernelGGL((k_gemm_wgrad<true>), grid, dim3(256), 0, s, N, kchunk, A, sa, B, sb, C, ldc, c_split,
                         E.colsum);
    else
      hipLaunchKernelGGL((k_gemm_wgrad<false>), grid, dim3(256), 0, s, N, kchunk, A, sa, B, sb, C, ldc, c_split,
                         nullptr);
  } else if (E.colsum) {
    if (b_kcontig) return hipErrorInvalidValue;
    if (a_kcontig) GLH_GEMM(true, false, true);
    else GLH_GEMM(false, false, true);
  } else if (a_kcontig && b_kcontig) GLH_GEMM(true, true, false);
  else if (a_kcontig) GLH_GEMM(true, false, false);
  else if (b_kcontig) GLH_GEMM(false, true, false);
  else GLH_GEMM(false, false, false);
#undef GLH_GEMM
  return hipGetLastError();
}

hipError_t grad_reduce(long long n_params, const GradLayout &L, const float *part, const float *cpart, float *G,
                       hipStream_t s)
{
  if (L.n_layers < 1 || L.n_layers > MLP_MAX_LAYERS) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_grad_reduce, dim3(grid1(n_params)), dim3(ET), 0, s, n_params, L, part, cpart, G);
  return hipGetLastError();
}

hipError_t adam_pack(long long n, float *p, const float *g, float *m, float *v, float lr, float b1, float b2, float eps,
                     const int *step, const GradLayout &L, bf16 *const *Wb, bf16 *const *WbT, hipStream_t s)
{
  if (L.n_layers < 1 || L.n_layers > MLP_MAX_LAYERS) return hipErrorInvalidValue;
  WbPtrs W{}, WT{};
  for (int l = 0; l < L.n_layers; ++l) {
    W.p[l] = Wb[l];
    WT.p[l] = WbT[l];
  }
  hipLaunchKernelGGL(k_adam_pack, dim3(grid1(n)), dim3(ET), 0, s, n, p, g, m, v, lr, b1, b2, eps, step, L, W, WT);
  return hipGetLastError();
}

hipError_t fourier(int n, int n_real, const float *X, const float *Bf, int F, bf16 *H, hipStream_t s)
{
  hipLaunchKernelGGL(k_fourier, dim3(grid1((long long)n * F)), dim3(ET), 0, s, n, n_real, X, Bf, F, H);
  return hipGetLastError();
}
hipError_t mse(int n, int n_real, int n_out, int ldy, const float *Y, const float *T, bf16 *dY, float *partial,
               hipStream_t s)
{
  hipLaunchKernelGGL(k_mse, dim3(grid1((long long)n * ldy)), dim3(ET), 0, s, n, n_real, n_out, ldy, Y, T, dY,
                     partial);
  return hipGetLastError();
}
int mse_blocks(int n, int ldy) { return (int)grid1((long long)n * ldy); }
hipError_t reduce1(int n, const float *in, float *out, hipStream_t s)
{
  hipLaunchKernelGGL(k_reduce1, dim3(1), dim3(ET), 0, s, n, in, out);
  return hipGetLastError();
}
hipError_t pack_w(int K, int N, int Np, const float *W, bf16 *Wb, bf16 *WbT, hipStream_t s)
{
  hipLaunchKernelGGL(k_pack_w, dim3(grid1((long long)K * Np)), dim3(ET), 0, s, K, N, Np, W, Wb, WbT);
  return hipGetLastError();
}
hipError_t gather(int nb, long long n_data, uint64_t seed, const int *step, const float *X, const float *Y, int n_out,
                  float *Xb, float *Yb, hipStream_t s)
{
  hipLaunchKernelGGL(k_gather, dim3(grid1(nb)), dim3(ET), 0, s, nb, n_data, seed, step, X, Y, n_out, Xb, Yb);
  return hipGetLastError();
}
hipError_t take_out(int n, int n_out, int ldy, const float *Yp, float *Y, hipStream_t s)
{
  hipLaunchKernelGGL(k_take_out, dim3(grid1((long long)n * n_out)), dim3(ET), 0, s, n, n_out, ldy, Yp, Y);
  return hipGetLastError();
}
hipError_t step_inc(int *step, hipStream_t s)
{
  hipLaunchKernelGGL(k_step_inc, dim3(1), dim3(64), 0, s, step);
  return hipGetLastError();
}

}  // namespace mlp
}  // namespace glh
