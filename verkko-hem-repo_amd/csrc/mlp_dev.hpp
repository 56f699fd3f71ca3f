// Launchers of the coordinate-MLP kernels (mlp.hip); plain device pointers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace glh {
namespace mlp {

struct bf16 {
  uint16_t bits;
};

// Epilogue of a GEMM (applied to the FP32 accumulators of C = A B):
//   GEMM_STORE      C (+)= A B (split z writes C + z*c_split); with colsum
//                   set (B not k-contiguous), the workgroups of row block 0
//                   also write the column sums of B over their k range to
//                   colsum[z*N + n] (the bias gradient of a layer whose
//                   weight gradient this is);
//   GEMM_BIAS_SILU  z = A B + bias[n]: Z[m*ldz + n] = z, H[m*ldh + n] = bf16(SiLU(z));
//   GEMM_BIAS       C = A B + bias[n] for n < nbias (A B beyond);
//   GEMM_SILU_BACK  H[m*ldh + n] = bf16((A B) SiLU'(Z[m*ldz + n])).
enum { GEMM_STORE = 0, GEMM_BIAS_SILU = 1, GEMM_BIAS = 2, GEMM_SILU_BACK = 3 };
struct GemmEpi {
  int mode = GEMM_STORE;
  const float *bias = nullptr;
  int nbias = 0;
  float *Z = nullptr;
  int ldz = 0;
  bf16 *H = nullptr;
  int ldh = 0;
  float *colsum = nullptr;
};

// C (FP32) = A B (+ C), A (M x K) and B (K x N) in BF16.  a_kcontig: A(m,k) =
// A[m*sa + k], else A[k*sa + m]; b_kcontig: B(k,n) = B[n*sb + k], else
// B[k*sb + n].  splits > 1: split K, split z writes C + z*c_split.
// M, N multiples of 64, K a multiple of 32*splits.
hipError_t gemm(int M, int N, int K, const bf16 *A, bool a_kcontig, long long sa, const bf16 *B, bool b_kcontig,
                long long sb, float *C, int ldc, bool accumulate, int splits, long long c_split, hipStream_t s,
                const GemmEpi &epi = GemmEpi{});

// Where the split-K partials of every layer's weight gradient (and the
// per-split column sums that are its bias gradient) sit, and where the flat
// gradient takes them: layer l has K[l] x N[l] weights (padded to Np[l]
// columns in the GEMMs) at woff[l] and N[l] biases at boff[l].
constexpr int MLP_MAX_LAYERS = 8;
struct GradLayout {
  int n_layers = 0, splits = 1;
  int K[MLP_MAX_LAYERS], N[MLP_MAX_LAYERS], Np[MLP_MAX_LAYERS];
  long long woff[MLP_MAX_LAYERS], boff[MLP_MAX_LAYERS], poff[MLP_MAX_LAYERS], coff[MLP_MAX_LAYERS];
};
// G[i] = the sum over the splits (in order) of parameter i's partials
hipError_t grad_reduce(long long n_params, const GradLayout &L, const float *part, const float *cpart, float *G,
                       hipStream_t s);
// Adam on the FP32 master parameters, then the BF16 weight copies of the
// updated weights: Wb[l] (K[l] x Np[l]) and WbT[l] (Np[l] x K[l])
hipError_t adam_pack(long long n, float *p, const float *g, float *m, float *v, float lr, float b1, float b2, float eps,
                     const int *step, const GradLayout &L, bf16 *const *Wb, bf16 *const *WbT, hipStream_t s);
// H[s, j] = sin(2 pi X[s].Bf[:, j]), H[s, F + j] = cos(...); rows >= n_real are 0
hipError_t fourier(int n, int n_real, const float *X, const float *Bf, int F, bf16 *H, hipStream_t s);
hipError_t mse(int n, int n_real, int n_out, int ldy, const float *Y, const float *T, bf16 *dY, float *partial,
               hipStream_t s);
int mse_blocks(int n, int ldy);
hipError_t reduce1(int n, const float *in, float *out, hipStream_t s);
// Wb (K x Np) and WbT (Np x K) = BF16(W) (K x N), zeros in the padding
hipError_t pack_w(int K, int N, int Np, const float *W, bf16 *Wb, bf16 *WbT, hipStream_t s);
hipError_t gather(int nb, long long n_data, uint64_t seed, const int *step, const float *X, const float *Y, int n_out,
                  float *Xb, float *Yb, hipStream_t s);
hipError_t take_out(int n, int n_out, int ldy, const float *Yp, float *Y, hipStream_t s);
hipError_t step_inc(int *step, hipStream_t s);

}  // namespace mlp
}  // namespace glh
