// C ABI of the MLP surrogate: parameters, fused BF16-MFMA inference,
// training (forward, backward, Adam) and the data generator.  Kernels are in
// surrogate.hpp; declarations in include/alya_sma.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <random>
#include <string>
#include <vector>

#include "alya_sma.h"
#include "common.hpp"
#include "surrogate.hpp"

namespace {
#define SCHK(x) SMA_HIPCHK(x)
unsigned nb(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }
}  // namespace

struct sma_mlp {
  // canonical FP32 master parameters and Adam state
  float *W = nullptr, *b = nullptr, *mW = nullptr, *vW = nullptr, *mb = nullptr, *vb = nullptr;
  float *gW = nullptr, *gb = nullptr;
  // padded copies read by the kernels
  __bf16* Wp = nullptr;
  __bf16* Wi = nullptr;        // inference layout (mlp::inf_index)
  float* bp = nullptr;
  int64_t step = 0;
  // training workspace (grown on demand)
  int64_t cap = 0;
  __bf16* X0 = nullptr;
  float* Z[5] = {};
  __bf16* Y[4] = {};
  float* C = nullptr;
  float* dZ = nullptr;
  __bf16* dZb = nullptr;
  float* part = nullptr;
  float* loss_part = nullptr;
  float* gWp = nullptr;
  float* gbp = nullptr;
  float* wloss = nullptr;
  std::vector<void*> ws;
};

namespace {

constexpr int KCHUNK = 512;

int grow(sma_mlp* m, int64_t M) {
  if (M <= m->cap) return 0;
  for (void* p : m->ws) (void)hipFree(p);
  m->ws.clear();
  auto A = [&](auto** p, size_t bytes) -> int {
    SCHK(hipMalloc((void**)p, bytes));
    m->ws.push_back((void*)*p);
    return 0;
  };
  int rc = 0;
  if ((rc = A(&m->X0, M * mlp::KP0 * sizeof(__bf16)))) return rc;
  for (int l = 0; l < 5; ++l)
    if ((rc = A(&m->Z[l], M * mlp::layer_n(l) * sizeof(float)))) return rc;
  for (int l = 0; l < 4; ++l)
    if ((rc = A(&m->Y[l], M * mlp::NH * sizeof(__bf16)))) return rc;
  if ((rc = A(&m->C, M * mlp::NH * sizeof(float)))) return rc;
  if ((rc = A(&m->dZ, M * mlp::NH * sizeof(float)))) return rc;
  if ((rc = A(&m->dZb, M * mlp::NH * sizeof(__bf16)))) return rc;
  const int64_t nz = (M + KCHUNK - 1) / KCHUNK;
  if ((rc = A(&m->part, nz * mlp::NH * mlp::NH * sizeof(float)))) return rc;
  if ((rc = A(&m->loss_part, nb(M * mlp::NP4, 256) * sizeof(float)))) return rc;
  m->cap = M;
  return 0;
}

void gemm(const __bf16* A, int64_t sam, int64_t sak, const __bf16* B, int64_t sbk, int64_t sbn,
          float* C, int M, int N, int K, int kchunk, hipStream_t s) {
  mlp::GemmArgs g{A, B, C, M, N, K, kchunk, sam, sak, sbk, sbn};
  const int waves = (M / 32) * (N / 32);
  const int nz = (K + kchunk - 1) / kchunk;
  hipLaunchKernelGGL(mlp::gemm_bf16, dim3(nb(waves, 4), 1, nz), dim3(256), 0, s, g);
}

// forward (training layout) + loss gradient + backward into gWp/gbp, then the
// canonical gradient (gW, gb).  Returns the mean loss through loss_out.
int forward_backward(sma_mlp* m, int64_t n, const float* X, const float* T, float* gW, float* gb,
                     float* loss_out, hipStream_t s) {
  const int M = (int)(((n + 31) / 32) * 32);
  int rc = grow(m, M);
  if (rc) return rc;
  hipLaunchKernelGGL(mlp::pack_input, dim3(nb((int64_t)M * mlp::KP0, 256)), dim3(256), 0, s, X, n, M, m->X0);
  const __bf16* in = m->X0;
  for (int l = 0; l < 5; ++l) {
    const int K = mlp::layer_k(l), N = mlp::layer_n(l);
    gemm(in, K, 1, m->Wp + mlp::woff(l), 1, K, m->C, M, N, K, K, s);
    hipLaunchKernelGGL(mlp::bias_act, dim3(nb((int64_t)M * N, 256)), dim3(256), 0, s, m->C,
                       m->bp + mlp::boff(l), M, N, m->Z[l], l < 4 ? m->Y[l] : nullptr, l < 4 ? 1 : 0);
    if (l < 4) in = m->Y[l];
  }
  const unsigned nlb = nb((int64_t)M * mlp::NP4, 256);
  hipLaunchKernelGGL(mlp::loss_grad, dim3(nlb), dim3(256), 0, s, m->Z[4], T, m->wloss, n, M, m->dZ,
                     m->dZb, m->loss_part);
  for (int l = 4; l >= 0; --l) {
    const int K = mlp::layer_k(l), N = mlp::layer_n(l);
    const __bf16* Xl = l == 0 ? m->X0 : m->Y[l - 1];
    const int nz = (M + KCHUNK - 1) / KCHUNK;
    // dW_l [N][K] = sum_m dZ[m][n] X_l[m][k]
    gemm(m->dZb, 1, N, Xl, K, 1, m->part, N, K, M, KCHUNK, s);
    hipLaunchKernelGGL(mlp::reduce_parts, dim3(nb((int64_t)N * K, 256)), dim3(256), 0, s, m->part,
                       m->gWp + mlp::woff(l), (int64_t)N * K, nz);
    // db_l [N] = sum_m dZ[m][n]
    hipLaunchKernelGGL(mlp::col_sums, dim3(nb(N, 128), nz), dim3(128), 0, s, m->dZ, M, N, KCHUNK, m->part);
    hipLaunchKernelGGL(mlp::reduce_parts, dim3(1), dim3(256), 0, s, m->part, m->gbp + mlp::boff(l),
                       (int64_t)N, nz);
    hipLaunchKernelGGL(mlp::unpad_grad, dim3(nb(mlp::NH * mlp::NH, 256)), dim3(256), 0, s,
                       m->gWp + mlp::woff(l), m->gbp + mlp::boff(l), gW, gb, l);
    if (l > 0) {
      // dY_{l-1} [M][K] = dZ [M][N] W_l [N][K]; dZ_{l-1} = dY * silu'(Z_{l-1})
      gemm(m->dZb, N, 1, m->Wp + mlp::woff(l), K, 1, m->C, M, K, N, N, s);
      hipLaunchKernelGGL(mlp::silu_back, dim3(nb((int64_t)M * K, 256)), dim3(256), 0, s, m->C,
                         m->Z[l - 1], (int64_t)M * K, m->dZ, m->dZb);
    }
  }
  SCHK(hipGetLastError());
  if (loss_out) {
    std::vector<float> lp(nlb);
    SCHK(hipMemcpyAsync(lp.data(), m->loss_part, nlb * sizeof(float), hipMemcpyDeviceToHost, s));
    SCHK(hipStreamSynchronize(s));
    double t = 0.0;
    for (float v : lp) t += v;
    *loss_out = (float)(t / (double)n);
  }
  return 0;
}

}  // namespace

extern "C" {

int sma_mlp_create(uint64_t seed, sma_mlp** out) {
  if (!out) return sma_abi::fail(SMA_ERR_ARG, "sma_mlp_create: null out");
  sma_mlp* m = new sma_mlp();
  auto A = [&](auto** p, size_t bytes) -> int {
    SCHK(hipMalloc((void**)p, bytes));
    SCHK(hipMemset(*p, 0, bytes));
    return 0;
  };
  int rc = 0;
  if ((rc = A(&m->W, mlp::W_TOTAL * 4)) || (rc = A(&m->b, mlp::B_TOTAL * 4)) ||
      (rc = A(&m->mW, mlp::W_TOTAL * 4)) || (rc = A(&m->vW, mlp::W_TOTAL * 4)) ||
      (rc = A(&m->mb, mlp::B_TOTAL * 4)) || (rc = A(&m->vb, mlp::B_TOTAL * 4)) ||
      (rc = A(&m->gW, mlp::W_TOTAL * 4)) || (rc = A(&m->gb, mlp::B_TOTAL * 4)) ||
      (rc = A(&m->Wp, mlp::WPAD_TOTAL * 2)) || (rc = A(&m->Wi, mlp::WLDS_TOTAL * 2)) ||
      (rc = A(&m->bp, mlp::BPAD_TOTAL * 4)) ||
      (rc = A(&m->gWp, mlp::WPAD_TOTAL * 4)) || (rc = A(&m->gbp, mlp::BPAD_TOTAL * 4)) ||
      (rc = A(&m->wloss, 16 * 4))) {
    sma_mlp_destroy(m);
    return rc;
  }
  // He-uniform initialisation, seeded
  std::mt19937_64 g(seed);
  std::vector<float> W(mlp::W_TOTAL), b(mlp::B_TOTAL, 0.f);
  for (int l = 0; l < 5; ++l) {
    const int kr = mlp::layer_kreal(l), nr = mlp::layer_nreal(l);
    const float lim = std::sqrt(6.0f / kr) * (l == 4 ? 0.5f : 1.0f);
    std::uniform_real_distribution<float> U(-lim, lim);
    for (int i = 0; i < nr * kr; ++i) W[mlp::woff_c(l) + i] = U(g);
  }
  rc = sma_mlp_set_params(m, W.data(), b.data());
  if (rc) { sma_mlp_destroy(m); return rc; }
  std::vector<float> w(16, 1.0f);
  if (hipMemcpy(m->wloss, w.data(), 16 * 4, hipMemcpyHostToDevice) != hipSuccess) {
    sma_mlp_destroy(m);
    return sma_abi::fail(SMA_ERR_HIP, "upload loss weights");
  }
  *out = m;
  return SMA_OK;
}

int sma_mlp_destroy(sma_mlp* m) {
  if (!m) return SMA_OK;
  for (void* p : m->ws) (void)hipFree(p);
  void* ps[] = {m->W, m->b, m->mW, m->vW, m->mb, m->vb, m->gW, m->gb, m->Wp, m->Wi, m->bp, m->gWp, m->gbp,
                m->wloss};
  for (void* p : ps) if (p) (void)hipFree(p);
  delete m;
  return SMA_OK;
}

int sma_mlp_set_params(sma_mlp* m, const float* W, const float* b) {
  if (!m || !W || !b) return sma_abi::fail(SMA_ERR_ARG, "sma_mlp_set_params: null argument");
  SCHK(hipMemcpy(m->W, W, mlp::W_TOTAL * 4, hipMemcpyHostToDevice));
  SCHK(hipMemcpy(m->b, b, mlp::B_TOTAL * 4, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(mlp::pack_params, dim3(nb(mlp::WPAD_TOTAL, 256)), dim3(256), 0, nullptr, m->W,
                     m->b, m->Wp, m->Wi, m->bp);
  SCHK(hipGetLastError());
  SCHK(hipDeviceSynchronize());
  return SMA_OK;
}

int sma_mlp_get_params(sma_mlp* m, float* W, float* b) {
  if (!m || !W || !b) return sma_abi::fail(SMA_ERR_ARG, "sma_mlp_get_params: null argument");
  SCHK(hipDeviceSynchronize());
  SCHK(hipMemcpy(W, m->W, mlp::W_TOTAL * 4, hipMemcpyDeviceToHost));
  SCHK(hipMemcpy(b, m->b, mlp::B_TOTAL * 4, hipMemcpyDeviceToHost));
  return SMA_OK;
}

int sma_mlp_set_loss_weights(sma_mlp* m, const float* w12) {
  if (!m || !w12) return sma_abi::fail(SMA_ERR_ARG, "null argument");
  std::vector<float> w(16, 0.f);
  std::memcpy(w.data(), w12, 12 * sizeof(float));
  SCHK(hipMemcpy(m->wloss, w.data(), 16 * 4, hipMemcpyHostToDevice));
  return SMA_OK;
}

int sma_mlp_forward_device(sma_mlp* m, int64_t n, const float* X, float* Y, void* stream) {
  if (!m || n < 0 || (n > 0 && (!X || !Y))) return sma_abi::fail(SMA_ERR_ARG, "sma_mlp_forward_device: bad arguments");
  if (n > INT32_MAX) return sma_abi::fail(SMA_ERR_ARG, "sma_mlp_forward_device: n must be below 2^31");
  if (n == 0) return SMA_OK;
  static bool attr = false;
  if (!attr) {
    SCHK(hipFuncSetAttribute((const void*)mlp::mlp_infer, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)mlp::INF_SMEM));
    attr = true;
  }
  const int64_t nblk = (n + mlp::INF_ROWS - 1) / mlp::INF_ROWS;
  int dev = 0, ncu = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  const unsigned grid = (unsigned)std::min<int64_t>(nblk, ncu);
  hipLaunchKernelGGL(mlp::mlp_infer, dim3(grid), dim3(mlp::INF_THREADS), mlp::INF_SMEM, (hipStream_t)stream, m->Wi,
                     m->bp, X, n, (const unsigned*)nullptr, (int64_t)0, Y);
  SCHK(hipGetLastError());
  return SMA_OK;
}

int sma_mlp_grad_device(sma_mlp* m, int64_t n, const float* X, const float* T, float* gW, float* gb,
                        float* loss_out, void* stream) {
  if (!m || n <= 0 || !X || !T || !gW || !gb) return sma_abi::fail(SMA_ERR_ARG, "sma_mlp_grad_device: bad arguments");
  return forward_backward(m, n, X, T, gW, gb, loss_out, (hipStream_t)stream);
}

int sma_mlp_train_step_device(sma_mlp* m, int64_t n, const float* X, const float* T, double lr,
                              float* loss_out, void* stream) {
  if (!m || n <= 0 || !X || !T) return sma_abi::fail(SMA_ERR_ARG, "sma_mlp_train_step_device: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  int rc = forward_backward(m, n, X, T, m->gW, m->gb, loss_out, s);
  if (rc) return rc;
  m->step += 1;
  mlp::AdamArgs A{m->W, m->b, m->mW, m->vW, m->mb, m->vb, m->gW, m->gb, m->Wp, m->bp, m->Wi,
                  (float)lr, 0.9f, 0.999f, 1e-8f,
                  (float)(1.0 - std::pow(0.9, (double)m->step)),
                  (float)(1.0 - std::pow(0.999, (double)m->step))};
  hipLaunchKernelGGL(mlp::adam_step, dim3(nb(mlp::W_TOTAL, 256)), dim3(256), 0, s, A);
  SCHK(hipGetLastError());
  return SMA_OK;
}

int sma_mlp_gen_samples_device(int64_t n, uint64_t seed, const double* props_dev, double strain_sd,
                               double step_sd, double jitter, float* X, float* T, uint32_t* flags,
                               void* stream) {
  if (n < 0 || (n > 0 && (!props_dev || !X || !T || !flags)))
    return sma_abi::fail(SMA_ERR_ARG, "sma_mlp_gen_samples_device: bad arguments");
  if (n == 0) return SMA_OK;
  mlp::GenArgs G{n, seed, props_dev, strain_sd, step_sd, jitter, X, T, flags};
  hipLaunchKernelGGL(mlp::gen_samples, dim3(nb(n, 256)), dim3(256), 0, (hipStream_t)stream, G);
  SCHK(hipGetLastError());
  return SMA_OK;
}

}  // extern "C"

// Internal: inference over the step's list of points (its segment counters
// are read on the device; seg_cap slots per segment).
int sma_mlp_infer_listed(sma_mlp* m, const float* X, const unsigned* seg_count, int64_t seg_cap,
                         float* Y, hipStream_t s) {
  static bool attr = false;
  if (!attr) {
    SCHK(hipFuncSetAttribute((const void*)mlp::mlp_infer, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)mlp::INF_SMEM));
    attr = true;
  }
  int dev = 0, ncu = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  const int64_t cap = seg_cap * mlp::LIST_SEGS;
  const unsigned grid =
      (unsigned)std::max<int64_t>(1, std::min<int64_t>((cap + mlp::INF_ROWS - 1) / mlp::INF_ROWS, ncu));
  hipLaunchKernelGGL(mlp::mlp_infer, dim3(grid), dim3(mlp::INF_THREADS), mlp::INF_SMEM, s, m->Wi, m->bp, X,
                     (int64_t)0, seg_count, seg_cap, Y);
  SCHK(hipGetLastError());
  return 0;
}
