// Launchers of the coordinate-MLP kernels (mlp.hip); plain device pointers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace glh {
namespace mlp {

struct bf16 {
  uint16_t bits;
};

// C (FP32) = A B (+ C), A (M x K) and B (K x N) in BF16.  a_kcontig: A(m,k) =
// A[m*sa + k], else A[k*sa + m]; b_kcontig: B(k,n) = B[n*sb + k], else
// B[k*sb + n].  splits > 1: split K, split z writes C + z*c_split.
// M, N multiples of 64, K a multiple of 32*splits.
hipError_t gemm(int M, int N, int K, const bf16 *A, bool a_kcontig, long long sa, const bf16 *B, bool b_kcontig,
                long long sb, float *C, int ldc, bool accumulate, int splits, long long c_split, hipStream_t s);
// H[s, j] = sin(2 pi X[s].Bf[:, j]), H[s, F + j] = cos(...); rows >= n_real are 0
hipError_t fourier(int n, int n_real, const float *X, const float *Bf, int F, bf16 *H, hipStream_t s);
hipError_t bias_silu(long long total, int N, float *Z, const float *bias, bf16 *H, hipStream_t s);
hipError_t bias_out(int n, int n_out, int ldy, float *Y, const float *bias, hipStream_t s);
hipError_t mse(int n, int n_real, int n_out, int ldy, const float *Y, const float *T, bf16 *dY, float *partial,
               hipStream_t s);
int mse_blocks(int n, int ldy);
hipError_t silu_back(long long total, const float *dH, const float *Z, bf16 *dZ, hipStream_t s);
hipError_t colsum(int n, int N, const bf16 *D, float *partial, int parts, float *out, hipStream_t s);
hipError_t sum_parts(int parts, long long len, const float *partial, float *out, hipStream_t s);
hipError_t reduce1(int n, const float *in, float *out, hipStream_t s);
hipError_t adam(long long n, float *p, const float *g, float *m, float *v, float lr, float b1, float b2, float eps,
                const int *step, hipStream_t s);
hipError_t pack_w(int K, int N, int Np, const float *W, bf16 *Wb, hipStream_t s);
hipError_t gather(int nb, long long n_data, uint64_t seed, const int *step, const float *X, const float *Y, int n_out,
                  float *Xb, float *Yb, hipStream_t s);
hipError_t unpad(int K, int N, int Np, const float *src, float *dst, hipStream_t s);
hipError_t take_out(int n, int n_out, int ldy, const float *Yp, float *Y, hipStream_t s);
hipError_t step_inc(int *step, hipStream_t s);

}  // namespace mlp
}  // namespace glh
