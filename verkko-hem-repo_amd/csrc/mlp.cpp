// Host side of the coordinate-MLP warm start (glh_mlp_* in include/glhip.h):
// parameter layout and initialisation, the training step (gather, forward,
// MSE, backward, Adam) replayed as a HIP graph, inference and the gradient
// entry used by the tests.
//
// Flat FP32 parameter layout, layer l = 0..depth: W_l (K_l x N_l, row-major,
// y = x W_l + b_l) then b_l (N_l).  K_0 = 2F (Fourier features), N_l = width
// for l < depth, N_depth = n_out.  The Fourier matrix (6 x F) is fixed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "glhip.h"
#include "mlp_dev.hpp"

using glh::mlp::bf16;

namespace {

// device buffer owned by the network handle; movable (vectors of layers)
template <typename T>
struct Buf {
  T *p = nullptr;
  size_t n = 0;
  Buf() = default;
  Buf(const Buf &) = delete;
  Buf &operator=(const Buf &) = delete;
  Buf(Buf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  Buf &operator=(Buf &&o) noexcept
  {
    if (this != &o) {
      release();
      p = o.p; n = o.n;
      o.p = nullptr; o.n = 0;
    }
    return *this;
  }
  ~Buf() { release(); }
  hipError_t alloc(size_t count)
  {
    if (count <= n && p) return hipSuccess;
    release();
    hipError_t e = hipMalloc((void **)&p, std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) { p = nullptr; return e; }
    n = count;
    return hipSuccess;
  }
  void release()
  {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

constexpr int OUTP = 64;       // padded width of the output layer
constexpr int DW_SPLITS = 64;  // K splits of the weight-gradient products

}  // namespace

struct glh_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int F = 64, width = 256, depth = 5, n_out = 18;
  std::vector<int> K, N, Np;             // per layer
  std::vector<long long> woff, boff;     // offsets in the flat parameters
  long long n_params = 0;
  Buf<float> P, G, Mo, Vo, Bf;
  std::vector<Buf<bf16>> Wb, WbT;  // BF16 weights, K x Np and transposed Np x K
  Buf<float> Xall, Yall;
  long long n_data = 0;
  // batch workspace
  int nb = 0;
  Buf<float> Xb, Yb, Yo, part, cpart, lpart, loss;
  std::vector<Buf<bf16>> H;
  std::vector<Buf<float>> Z;
  Buf<bf16> dY, dZa, dZb;
  Buf<int> step;
  hipGraphExec_t graph = nullptr;
  int graph_nb = 0;
  uint64_t graph_seed = 0;
  float graph_lr = 0.f;
};

namespace {

int fail(glh_mlp *m, int code, const std::string &msg)
{
  if (m) m->err = msg;
  return code;
}

#define MCHK(m, ...)                                                                              \
  do {                                                                                            \
    hipError_t e_ = (__VA_ARGS__);                                                                \
    if (e_ != hipSuccess) return fail((m), GLH_ERR_HIP, std::string(#__VA_ARGS__) + ": " + hipGetErrorString(e_)); \
  } while (0)
#define RCHK(expr)       \
  do {                   \
    int rc_ = (expr);    \
    if (rc_) return rc_; \
  } while (0)

int pack_all(glh_mlp *m)
{
  for (int l = 0; l <= m->depth; ++l)
    MCHK(m, glh::mlp::pack_w(m->K[l], m->N[l], m->Np[l], m->P.p + m->woff[l], m->Wb[l].p, m->WbT[l].p,
                             m->stream));
  return GLH_OK;
}

int ensure_batch(glh_mlp *m, int nb)
{
  if (nb % 64 || nb < 64) return fail(m, GLH_ERR_ARG, "batch must be a positive multiple of 64");
  if (m->nb == nb) return GLH_OK;
  m->H.resize(m->depth + 1);
  m->Z.resize(m->depth);
  MCHK(m, m->Xb.alloc((size_t)nb * 6));
  MCHK(m, m->Yb.alloc((size_t)nb * m->n_out));
  MCHK(m, m->H[0].alloc((size_t)nb * 2 * m->F));
  for (int l = 0; l < m->depth; ++l) {
    MCHK(m, m->H[l + 1].alloc((size_t)nb * m->width));
    MCHK(m, m->Z[l].alloc((size_t)nb * m->width));
  }
  MCHK(m, m->Yo.alloc((size_t)nb * OUTP));
  MCHK(m, m->dY.alloc((size_t)nb * OUTP));
  MCHK(m, m->dZa.alloc((size_t)nb * m->width));
  MCHK(m, m->dZb.alloc((size_t)nb * m->width));
  // split-K partials of every layer's weight and bias gradients at once (the
  // gradient is summed from them in one pass after the backward sweep)
  size_t ptot = 0, ctot = 0;
  for (int l = 0; l <= m->depth; ++l) {
    ptot += (size_t)m->K[l] * m->Np[l];
    ctot += (size_t)m->Np[l];
  }
  MCHK(m, m->part.alloc(ptot * DW_SPLITS));
  MCHK(m, m->cpart.alloc(ctot * DW_SPLITS));
  MCHK(m, m->lpart.alloc((size_t)glh::mlp::mse_blocks(nb, OUTP)));
  MCHK(m, m->loss.alloc(1));
  m->nb = nb;
  if (m->graph) (void)hipGraphExecDestroy(m->graph);
  m->graph = nullptr;
  return GLH_OK;
}

// weight-gradient splits of a batch of nb rows (fixed slices of the batch)
int dw_splits(int nb)
{
  int sp = 1;
  while (sp < DW_SPLITS && (nb / 32) % (2 * sp) == 0) sp *= 2;
  return sp;
}

glh::mlp::GradLayout grad_layout(const glh_mlp *m, int sp)
{
  glh::mlp::GradLayout L;
  L.n_layers = m->depth + 1;
  L.splits = sp;
  long long po = 0, co = 0;
  for (int l = 0; l <= m->depth; ++l) {
    L.K[l] = m->K[l]; L.N[l] = m->N[l]; L.Np[l] = m->Np[l];
    L.woff[l] = m->woff[l]; L.boff[l] = m->boff[l];
    L.poff[l] = po; L.coff[l] = co;
    po += (long long)sp * m->K[l] * m->Np[l];
    co += (long long)sp * m->Np[l];
  }
  return L;
}

// forward over the nb rows in Xb (the first n_real are real points); with
// grad, also the MSE against Yb and the full gradient into G.  The bias and
// SiLU of every hidden layer run in its GEMM's epilogue, the SiLU derivative
// in the epilogue of the input-gradient GEMM, and the bias gradients beside
// the weight gradients in their GEMMs (column sums of dZ).
int forward_backward(glh_mlp *m, int n_real, bool grad)
{
  using glh::mlp::GemmEpi;
  const int nb = m->nb, W = m->width, D = m->depth;
  hipStream_t s = m->stream;
  MCHK(m, glh::mlp::fourier(nb, n_real, m->Xb.p, m->Bf.p, m->F, m->H[0].p, s));
  for (int l = 0; l < D; ++l) {
    GemmEpi E;
    E.mode = glh::mlp::GEMM_BIAS_SILU;
    E.bias = m->P.p + m->boff[l];
    E.Z = m->Z[l].p;
    E.ldz = W;
    E.H = m->H[l + 1].p;
    E.ldh = W;
    MCHK(m, glh::mlp::gemm(nb, W, m->K[l], m->H[l].p, true, m->K[l], m->WbT[l].p, true, m->K[l], m->Z[l].p, W,
                           false, 1, 0, s, E));
  }
  {
    GemmEpi E;
    E.mode = glh::mlp::GEMM_BIAS;
    E.bias = m->P.p + m->boff[D];
    E.nbias = m->n_out;
    MCHK(m, glh::mlp::gemm(nb, OUTP, m->K[D], m->H[D].p, true, m->K[D], m->WbT[D].p, true, m->K[D], m->Yo.p, OUTP,
                           false, 1, 0, s, E));
  }
  if (!grad) return GLH_OK;
  const int nlp = glh::mlp::mse_blocks(nb, OUTP);
  MCHK(m, glh::mlp::mse(nb, n_real, m->n_out, OUTP, m->Yo.p, m->Yb.p, m->dY.p, m->lpart.p, s));
  MCHK(m, glh::mlp::reduce1(nlp, m->lpart.p, m->loss.p, s));
  // backward: dZ of layer l is dY (l = D) or a BF16 buffer alternating A/B.
  // The weight gradients split the batch into sp fixed slices.
  const int sp = dw_splits(nb);
  const glh::mlp::GradLayout L = grad_layout(m, sp);
  const bf16 *dZ = m->dY.p;
  int ldz = OUTP;
  for (int l = D; l >= 0; --l) {
    const int Kl = m->K[l], Npl = m->Np[l];
    // dW_l = H_l^T dZ_l per slice, and the slice's column sums of dZ_l (db_l)
    GemmEpi Ew;
    Ew.colsum = m->cpart.p + L.coff[l];
    MCHK(m, glh::mlp::gemm(Kl, Npl, nb, m->H[l].p, false, Kl, dZ, false, ldz, m->part.p + L.poff[l], Npl, false, sp,
                           (long long)Kl * Npl, s, Ew));
    if (l == 0) break;
    // dZ_{l-1} = (dZ_l W_l^T) * SiLU'(Z_{l-1})
    bf16 *dZn = (dZ == m->dZa.p) ? m->dZb.p : m->dZa.p;
    GemmEpi Eb;
    Eb.mode = glh::mlp::GEMM_SILU_BACK;
    Eb.Z = m->Z[l - 1].p;
    Eb.ldz = Kl;
    Eb.H = dZn;
    Eb.ldh = Kl;
    MCHK(m, glh::mlp::gemm(nb, Kl, Npl, dZ, true, ldz, m->Wb[l].p, true, Npl, nullptr, Kl, false, 1, 0, s, Eb));
    dZ = dZn;
    ldz = Kl;
  }
  // the flat gradient: every parameter's slices summed in order
  MCHK(m, glh::mlp::grad_reduce(m->n_params, L, m->part.p, m->cpart.p, m->G.p, s));
  return GLH_OK;
}

int train_step(glh_mlp *m, uint64_t seed, float lr)
{
  hipStream_t s = m->stream;
  MCHK(m, glh::mlp::step_inc(m->step.p, s));
  MCHK(m, glh::mlp::gather(m->nb, m->n_data, seed, m->step.p, m->Xall.p, m->Yall.p, m->n_out, m->Xb.p, m->Yb.p, s));
  RCHK(forward_backward(m, m->nb, true));
  // Adam, and the BF16 copies of the updated weights for the next step
  std::vector<bf16 *> wb(m->depth + 1), wbt(m->depth + 1);
  for (int l = 0; l <= m->depth; ++l) {
    wb[l] = m->Wb[l].p;
    wbt[l] = m->WbT[l].p;
  }
  MCHK(m, glh::mlp::adam_pack(m->n_params, m->P.p, m->G.p, m->Mo.p, m->Vo.p, lr, 0.9f, 0.999f, 1e-8f, m->step.p,
                              grad_layout(m, 1), wb.data(), wbt.data(), s));
  return GLH_OK;
}

}  // namespace

extern "C" {

int glh_mlp_create(int device, int n_freq, double sigma, int width, int depth, int n_out, uint64_t seed,
                   glh_mlp **out)
{
  if (!out) return GLH_ERR_ARG;
  *out = nullptr;
  if (n_freq < 1 || width < 64 || width % 64 || depth < 1 || depth + 1 > glh::mlp::MLP_MAX_LAYERS || n_out < 1 ||
      n_out > OUTP || (2 * n_freq) % 64 ||
      !(sigma > 0))
    return GLH_ERR_ARG;
  auto *m = new glh_mlp();
  m->device = device;
  m->F = n_freq; m->width = width; m->depth = depth; m->n_out = n_out;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) {
    delete m;
    return GLH_ERR_HIP;
  }
  for (int l = 0; l <= depth; ++l) {
    m->K.push_back(l == 0 ? 2 * n_freq : width);
    m->N.push_back(l == depth ? n_out : width);
    m->Np.push_back(l == depth ? OUTP : width);
  }
  for (int l = 0; l <= depth; ++l) {
    m->woff.push_back(m->n_params);
    m->n_params += (long long)m->K[l] * m->N[l];
    m->boff.push_back(m->n_params);
    m->n_params += m->N[l];
  }
  // U(-1/sqrt(K), 1/sqrt(K)) weights and biases; Fourier matrix N(0, sigma^2)
  std::mt19937_64 gen(seed);
  auto uni = [&]() { return (double)(gen() >> 11) * (1.0 / 9007199254740992.0); };
  std::vector<float> P(m->n_params);
  for (int l = 0; l <= depth; ++l) {
    const double a = 1.0 / std::sqrt((double)m->K[l]);
    for (long long i = m->woff[l]; i < m->boff[l] + m->N[l]; ++i) P[i] = (float)(a * (2.0 * uni() - 1.0));
  }
  std::vector<float> B(6 * n_freq);
  for (auto &b : B) {
    const double u1 = std::max(uni(), 1e-300), u2 = uni();
    b = (float)(sigma * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
  }
  int rc = GLH_OK;
  auto chk = [&](hipError_t e) {
    if (e != hipSuccess && rc == GLH_OK) rc = GLH_ERR_HIP;
  };
  chk(m->P.alloc(m->n_params));
  chk(m->G.alloc(m->n_params));
  chk(m->Mo.alloc(m->n_params));
  chk(m->Vo.alloc(m->n_params));
  chk(m->Bf.alloc(B.size()));
  chk(m->step.alloc(1));
  m->Wb.resize(depth + 1);
  m->WbT.resize(depth + 1);
  for (int l = 0; l <= depth; ++l) {
    chk(m->Wb[l].alloc((size_t)m->K[l] * m->Np[l]));
    chk(m->WbT[l].alloc((size_t)m->K[l] * m->Np[l]));
  }
  if (rc == GLH_OK) {
    chk(hipMemcpyAsync(m->P.p, P.data(), P.size() * sizeof(float), hipMemcpyHostToDevice, m->stream));
    chk(hipMemcpyAsync(m->Bf.p, B.data(), B.size() * sizeof(float), hipMemcpyHostToDevice, m->stream));
    chk(hipMemsetAsync(m->G.p, 0, m->n_params * sizeof(float), m->stream));
    chk(hipMemsetAsync(m->Mo.p, 0, m->n_params * sizeof(float), m->stream));
    chk(hipMemsetAsync(m->Vo.p, 0, m->n_params * sizeof(float), m->stream));
    chk(hipMemsetAsync(m->step.p, 0, sizeof(int), m->stream));
    if (rc == GLH_OK) rc = pack_all(m);
    chk(hipStreamSynchronize(m->stream));
  }
  if (rc != GLH_OK) {
    glh_mlp_destroy(m);
    return rc;
  }
  *out = m;
  return GLH_OK;
}

int glh_mlp_destroy(glh_mlp *m)
{
  if (!m) return GLH_OK;
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  if (m->graph) (void)hipGraphExecDestroy(m->graph);
  Buf<float> *fb[] = {&m->P, &m->G, &m->Mo, &m->Vo, &m->Bf, &m->Xall, &m->Yall, &m->Xb, &m->Yb,
                      &m->Yo, &m->part, &m->cpart, &m->lpart, &m->loss};
  for (auto *b : fb) b->release();
  for (auto &b : m->Wb) b.release();
  for (auto &b : m->WbT) b.release();
  for (auto &b : m->H) b.release();
  for (auto &b : m->Z) b.release();
  m->dY.release();
  m->dZa.release();
  m->dZb.release();
  m->step.release();
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return GLH_OK;
}

const char *glh_mlp_last_error(const glh_mlp *m) { return m ? m->err.c_str() : "null handle"; }

int64_t glh_mlp_n_params(const glh_mlp *m) { return m ? m->n_params : -1; }

int glh_mlp_get_params(glh_mlp *m, float *params, float *fourier)
{
  if (!m) return GLH_ERR_ARG;
  if (params) MCHK(m, hipMemcpyAsync(params, m->P.p, m->n_params * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  if (fourier) MCHK(m, hipMemcpyAsync(fourier, m->Bf.p, 6 * m->F * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MCHK(m, hipStreamSynchronize(m->stream));
  return GLH_OK;
}

int glh_mlp_set_params(glh_mlp *m, const float *params)
{
  if (!m || !params) return GLH_ERR_ARG;
  MCHK(m, hipMemcpyAsync(m->P.p, params, m->n_params * sizeof(float), hipMemcpyHostToDevice, m->stream));
  RCHK(pack_all(m));
  MCHK(m, hipStreamSynchronize(m->stream));
  return GLH_OK;
}

int glh_mlp_set_data(glh_mlp *m, int64_t n, const float *X, const float *Y)
{
  if (!m || n < 1 || !X || !Y) return GLH_ERR_ARG;
  MCHK(m, m->Xall.alloc((size_t)n * 6));
  MCHK(m, m->Yall.alloc((size_t)n * m->n_out));
  MCHK(m, hipMemcpyAsync(m->Xall.p, X, (size_t)n * 6 * sizeof(float), hipMemcpyHostToDevice, m->stream));
  MCHK(m, hipMemcpyAsync(m->Yall.p, Y, (size_t)n * m->n_out * sizeof(float), hipMemcpyHostToDevice, m->stream));
  MCHK(m, hipStreamSynchronize(m->stream));
  m->n_data = n;
  if (m->graph) (void)hipGraphExecDestroy(m->graph);
  m->graph = nullptr;
  return GLH_OK;
}

int glh_mlp_train(glh_mlp *m, int steps, int batch, double lr, uint64_t seed, int loss_every, float *loss_out)
{
  if (!m || steps < 0 || !(lr > 0)) return GLH_ERR_ARG;
  if (m->n_data < 1) return fail(m, GLH_ERR_STATE, "no training data");
  RCHK(ensure_batch(m, batch));
  hipStream_t s = m->stream;
  if (!m->graph || m->graph_seed != seed || m->graph_lr != (float)lr) {
    if (m->graph) (void)hipGraphExecDestroy(m->graph);
    m->graph = nullptr;
    hipGraph_t g = nullptr;
    MCHK(m, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int rc = train_step(m, seed, (float)lr);
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    MCHK(m, e);
    e = hipGraphInstantiate(&m->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    MCHK(m, e);
    m->graph_seed = seed;
    m->graph_lr = (float)lr;
  }
  int k_out = 0;
  for (int it = 0; it < steps; ++it) {
    MCHK(m, hipGraphLaunch(m->graph, s));
    if (loss_out && loss_every > 0 && (it + 1) % loss_every == 0) {
      float L = 0.f;
      MCHK(m, hipMemcpyAsync(&L, m->loss.p, sizeof(float), hipMemcpyDeviceToHost, s));
      MCHK(m, hipStreamSynchronize(s));
      loss_out[k_out++] = L / ((float)batch * (float)m->n_out);
    }
  }
  MCHK(m, hipStreamSynchronize(s));
  return GLH_OK;
}

int glh_mlp_loss_grad(glh_mlp *m, int n, const float *X, const float *Y, float *loss, float *grad)
{
  if (!m || n < 1 || !X || !Y) return GLH_ERR_ARG;
  const int nb = (n + 63) / 64 * 64;
  RCHK(ensure_batch(m, nb));
  hipStream_t s = m->stream;
  MCHK(m, hipMemsetAsync(m->Xb.p, 0, (size_t)nb * 6 * sizeof(float), s));
  MCHK(m, hipMemsetAsync(m->Yb.p, 0, (size_t)nb * m->n_out * sizeof(float), s));
  MCHK(m, hipMemcpyAsync(m->Xb.p, X, (size_t)n * 6 * sizeof(float), hipMemcpyHostToDevice, s));
  MCHK(m, hipMemcpyAsync(m->Yb.p, Y, (size_t)n * m->n_out * sizeof(float), hipMemcpyHostToDevice, s));
  RCHK(forward_backward(m, n, true));
  float L = 0.f;
  MCHK(m, hipMemcpyAsync(&L, m->loss.p, sizeof(float), hipMemcpyDeviceToHost, s));
  if (grad) MCHK(m, hipMemcpyAsync(grad, m->G.p, m->n_params * sizeof(float), hipMemcpyDeviceToHost, s));
  MCHK(m, hipStreamSynchronize(s));
  if (loss) *loss = L / ((float)n * (float)m->n_out);
  return GLH_OK;
}

int glh_mlp_predict(glh_mlp *m, int64_t n, const float *X, float *Y)
{
  if (!m || n < 1 || !X || !Y) return GLH_ERR_ARG;
  hipStream_t s = m->stream;
  // in chunks of at most 65536 points
  const int chunk = 65536;
  Buf<float> Yt;
  MCHK(m, Yt.alloc((size_t)chunk * m->n_out));
  for (int64_t i0 = 0; i0 < n; i0 += chunk) {
    const int nr = (int)std::min<int64_t>(chunk, n - i0);
    const int nb = (nr + 63) / 64 * 64;
    if (m->nb < nb) {
      const int rc = ensure_batch(m, nb);
      if (rc) { Yt.release(); return rc; }
    }
    hipError_t e = hipMemsetAsync(m->Xb.p, 0, (size_t)m->nb * 6 * sizeof(float), s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(m->Xb.p, X + i0 * 6, (size_t)nr * 6 * sizeof(float), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) { Yt.release(); return fail(m, GLH_ERR_HIP, hipGetErrorString(e)); }
    // forward over the allocated batch rows (rows beyond nr are zero inputs)
    const int rc = forward_backward(m, nr, false);
    if (rc) { Yt.release(); return rc; }
    e = glh::mlp::take_out(nr, m->n_out, OUTP, m->Yo.p, Yt.p, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(Y + i0 * m->n_out, Yt.p, (size_t)nr * m->n_out * sizeof(float), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) { Yt.release(); return fail(m, GLH_ERR_HIP, hipGetErrorString(e)); }
  }
  Yt.release();
  return GLH_OK;
}

}  // extern "C"
