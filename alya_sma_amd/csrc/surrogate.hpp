// MLP surrogate of the SMA return map on gfx950 BF16 MFMA.
//
// Network 16 -> 128 -> 128 -> 128 -> 128 -> 12, SiLU on the hidden layers,
// BF16 weights and activations, FP32 accumulation (v_mfma_f32_16x16x32_bf16).
// Device layouts pad the input to 32 features and the output to 32 so every
// GEMM is made of 16x16x32 tiles:
//   layer 0: W [128][32]   layers 1-3: W [128][128]   layer 4: W [32][128]
// (row n holds the weights into output n; rows/cols beyond the real sizes
// are zero).  The reference has no surrogate; its closest analogue is
// ann_framework%eval(nvec, input, output) (alya-ann/mod_ann_framework.f90:
// 489-556), which scales inputs, runs the network and unscales outputs per
// Gauss point, as sma_surrogate_eval does here.
//
// MFMA 16x16x32 lane layout (wave64), lane l:
//   A fragment: A[row = l%16][k = 8*(l/16) + j], j = 0..7
//   B fragment: B[k = 8*(l/16) + j][col = l%16]
//   D fragment: D[row = 4*(l/16) + i][col = l%16], i = 0..3
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "sma_law.hpp"
#include "surrogate_features.hpp"

namespace mlp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

constexpr int KP0 = 32, NP4 = 32;      // padded input / output widths
constexpr int NLAYER = 5;
__host__ __device__ constexpr int layer_k(int l) { return l == 0 ? KP0 : NH; }
__host__ __device__ constexpr int layer_n(int l) { return l == 4 ? NP4 : NH; }
__host__ __device__ constexpr int layer_kreal(int l) { return l == 0 ? NIN : NH; }
__host__ __device__ constexpr int layer_nreal(int l) { return l == 4 ? NOUT : NH; }
// offsets (elements) of each padded layer in the packed weight / bias arrays
__host__ __device__ constexpr int woff(int l) {
  return l == 0 ? 0 : l == 1 ? NH * KP0 : l == 2 ? NH * KP0 + NH * NH
       : l == 3 ? NH * KP0 + 2 * NH * NH : NH * KP0 + 3 * NH * NH;
}
constexpr int WPAD_TOTAL = NH * KP0 + 3 * NH * NH + NP4 * NH;   // 57344
__host__ __device__ constexpr int boff(int l) { return l * NH; }
constexpr int BPAD_TOTAL = 4 * NH + NP4;                         // 544
// canonical (unpadded) master layout: W_l [nreal][kreal] then all biases
__host__ __device__ constexpr int woff_c(int l) {
  return l == 0 ? 0 : l == 1 ? NH * NIN : l == 2 ? NH * NIN + NH * NH
       : l == 3 ? NH * NIN + 2 * NH * NH : NH * NIN + 3 * NH * NH;
}
constexpr int W_TOTAL = NH * NIN + 3 * NH * NH + NOUT * NH;      // 52736
constexpr int B_TOTAL = 4 * NH + NOUT;                           // 524

// output rows of a layer held in LDS (the output layer needs its first 16)
__host__ __device__ constexpr int inf_n(int l) { return l == NLAYER - 1 ? 16 : NH; }
static_assert(NOUT <= 16, "the output layer is one 16-feature block");
// Inference layout of the weights (a device copy, kept by pack_params and
// adam_step, that the inference kernel copies into LDS as it is).  Rows of
// K + 8 BF16 (16 more bytes) per row.  An A fragment is
// 16 rows x 16 bytes read at once; with rows of K = 128 (or 32) BF16 every
// row starts in the same LDS bank group, and the 16 reads would queue on it.
__host__ __device__ constexpr int wld(int l) { return layer_k(l) + 8; }
__host__ __device__ constexpr int woff_lds(int l) {
  return l == 0 ? 0 : woff_lds(l - 1) + inf_n(l - 1) * wld(l - 1);
}
constexpr int WLDS_TOTAL = woff_lds(NLAYER);
constexpr int BLDS_TOTAL = (NLAYER - 1) * NH + 16;
// input position, in a hidden layer's stored weight row, of the previous
// layer's feature f = 16 b + 4 q + i: 32 (b >> 1) + 8 q + 4 (b & 1) + i
__host__ __device__ constexpr int kperm(int f) {
  return 32 * (f >> 5) + 8 * ((f >> 2) & 3) + 4 * ((f >> 4) & 1) + (f & 3);
}

constexpr bool kperm_is_bijection() {
  bool seen[NH] = {};
  for (int f = 0; f < NH; ++f) {
    const int p = kperm(f);
    if (p < 0 || p >= NH || seen[p]) return false;
    seen[p] = true;
  }
  return true;
}
static_assert(kperm_is_bijection(), "kperm permutes the hidden inputs");

// index of padded weight (l, n, k) in the inference layout; -1: not kept
__host__ __device__ constexpr int inf_index(int l, int n, int k) {
  return n < inf_n(l) ? woff_lds(l) + n * wld(l) + (l == 0 ? k : kperm(k)) : -1;
}


// z sigmoid(z) with the hardware reciprocal (1 ulp): the result is rounded to
// BF16 (8 significant bits) before it is used
__device__ __forceinline__ float silu(float z) {
  return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}
__device__ __forceinline__ float silu_grad(float z) {
  const float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

// ----------------------------------------------------------------------------
// Generic BF16 GEMM on MFMA: C[z][M][N] = sum_{k in chunk z} A[m][k] B[k][n]
// A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn].  M, N multiples of 32,
// every K chunk a multiple of 32.  One wave computes one 32x32 block.
// ----------------------------------------------------------------------------
struct GemmArgs {
  const __bf16* A;
  const __bf16* B;
  float* C;
  int M, N, K, kchunk;
  int64_t sam, sak, sbk, sbn;
};

__device__ __forceinline__ bf16x8 load8(const __bf16* p, int64_t stride) {
  bf16x8 v;
  if (stride == 1) {
    v = *reinterpret_cast<const bf16x8*>(p);
  } else {
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = p[j * stride];
  }
  return v;
}

__global__ void __launch_bounds__(256) gemm_bf16(GemmArgs g) {
  const int lane = threadIdx.x & 63;
  const int wave = (blockIdx.x * 4) + (threadIdx.x >> 6);
  const int ntn = g.N / 32;
  const int tm = wave / ntn, tn = wave % ntn;
  if (tm * 32 >= g.M) return;
  const int m0 = tm * 32, n0 = tn * 32;
  const int z = blockIdx.z;
  const int kbeg = z * g.kchunk, kend = min(g.K, kbeg + g.kchunk);
  f32x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int r = lane & 15, kq = 8 * (lane >> 4);
  for (int k0 = kbeg; k0 < kend; k0 += 32) {
    bf16x8 af[2], bfr[2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
      af[a] = load8(g.A + (int64_t)(m0 + 16 * a + r) * g.sam + (int64_t)(k0 + kq) * g.sak, g.sak);
#pragma unroll
    for (int b = 0; b < 2; ++b)
      bfr[b] = load8(g.B + (int64_t)(k0 + kq) * g.sbk + (int64_t)(n0 + 16 * b + r) * g.sbn, g.sbk);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
  }
  float* C = g.C + (int64_t)z * g.M * g.N;
  const int rq = 4 * (lane >> 4);
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        C[(int64_t)(m0 + 16 * a + rq + i) * g.N + n0 + 16 * b + r] = acc[a][b][i];
}

// Sum of split-K partials in a fixed order (deterministic).
__global__ void reduce_parts(const float* __restrict__ part, float* __restrict__ out, int64_t n, int nz) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int z = 0; z < nz; ++z) s += part[(int64_t)z * n + i];
  out[i] = s;
}

// Z = C + b; hidden layers: Y = bf16(silu(Z)); output layer: Yout = Z (fp32)
__global__ void bias_act(const float* __restrict__ C, const float* __restrict__ b, int M, int N,
                         float* __restrict__ Z, __bf16* __restrict__ Y, int act) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)M * N) return;
  const int n = i % N;
  const float z = C[i] + b[n];
  Z[i] = z;
  if (act) Y[i] = (__bf16)silu(z);
}

// Pack fp32 features [M][16] (rows >= n zero) into bf16 [M][32].
__global__ void pack_input(const float* __restrict__ X, int64_t n, int M, __bf16* __restrict__ X0) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)M * KP0) return;
  const int64_t m = i / KP0;
  const int k = i % KP0;
  X0[i] = (__bf16)((m < n && k < NIN) ? X[m * NIN + k] : 0.f);
}

// Output-layer gradient of L = (1/n) sum_m sum_j w_j (y - t)^2 and the
// per-block partial losses.
__global__ void __launch_bounds__(256) loss_grad(const float* __restrict__ Z4, const float* __restrict__ T,
                                                 const float* __restrict__ w, int64_t n, int M,
                                                 float* __restrict__ dZ, __bf16* __restrict__ dZb,
                                                 float* __restrict__ loss_part) {
  __shared__ float red[256];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (i < (int64_t)M * NP4) {
    const int64_t m = i / NP4;
    const int j = i % NP4;
    float g = 0.f;
    if (m < n && j < NOUT) {
      const float d = Z4[i] - T[m * NOUT + j];
      l = w[j] * d * d;
      g = 2.f * w[j] * d / (float)n;
    }
    dZ[i] = g;
    dZb[i] = (__bf16)g;
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

// dZ = dY * silu'(Z)
__global__ void silu_back(const float* __restrict__ dY, const float* __restrict__ Z, int64_t count,
                          float* __restrict__ dZ, __bf16* __restrict__ dZb) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const float g = dY[i] * silu_grad(Z[i]);
  dZ[i] = g;
  dZb[i] = (__bf16)g;
}

// Column sums of dZ [M][N] over row chunks: part[chunk][n] (deterministic).
__global__ void col_sums(const float* __restrict__ dZ, int M, int N, int rows_per_chunk,
                         float* __restrict__ part) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (n >= N) return;
  const int r0 = c * rows_per_chunk, r1 = min(M, r0 + rows_per_chunk);
  float s = 0.f;
  for (int m = r0; m < r1; ++m) s += dZ[(int64_t)m * N + n];
  part[(int64_t)c * N + n] = s;
}

// Padded gradient (layer l) -> canonical gradient layout.
__global__ void unpad_grad(const float* __restrict__ gWp, const float* __restrict__ gbp,
                           float* __restrict__ gW, float* __restrict__ gb, int l) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int K = layer_k(l), kr = layer_kreal(l), nr = layer_nreal(l);
  if (i < nr * kr) {
    const int n = i / kr, k = i % kr;
    gW[woff_c(l) + i] = gWp[n * K + k];
  }
  if (i < nr) gb[boff(l) + i] = gbp[i];
}

// Adam on the canonical FP32 master parameters; refreshes the padded BF16
// weights and FP32 biases the kernels read.
struct AdamArgs {
  float* W; float* b;            // master, canonical
  float* mW; float* vW; float* mb; float* vb;
  const float* gW; const float* gb;
  __bf16* Wp; float* bp;         // padded device copies
  __bf16* Wi;                    // inference layout
  float lr, beta1, beta2, eps, bc1, bc2;
};

__device__ __forceinline__ void adam1(float& p, float& m, float& v, float g, const AdamArgs& A) {
  m = A.beta1 * m + (1.f - A.beta1) * g;
  v = A.beta2 * v + (1.f - A.beta2) * g * g;
  const float mh = m / A.bc1, vh = v / A.bc2;
  p -= A.lr * mh / (sqrtf(vh) + A.eps);
}

__global__ void adam_step(AdamArgs A) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < W_TOTAL) {
    adam1(A.W[i], A.mW[i], A.vW[i], A.gW[i], A);
    int l = 0;
    while (l < 4 && i >= woff_c(l + 1)) ++l;
    const int kr = layer_kreal(l), K = layer_k(l);
    const int j = i - woff_c(l), n = j / kr, k = j % kr;
    A.Wp[woff(l) + n * K + k] = (__bf16)A.W[i];
    const int ii = inf_index(l, n, k);
    if (ii >= 0) A.Wi[ii] = (__bf16)A.W[i];
  }
  if (i < B_TOTAL) {
    adam1(A.b[i], A.mb[i], A.vb[i], A.gb[i], A);
    A.bp[i] = A.b[i];
  }
}

// Canonical FP32 params -> padded device copies (pads zero).
__global__ void pack_params(const float* __restrict__ W, const float* __restrict__ b,
                            __bf16* __restrict__ Wp, __bf16* __restrict__ Wi, float* __restrict__ bp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < WPAD_TOTAL) {
    int l = 0;
    while (l < 4 && i >= woff(l + 1)) ++l;
    const int K = layer_k(l), j = i - woff(l), n = j / K, k = j % K;
    float v = 0.f;
    if (n < layer_nreal(l) && k < layer_kreal(l)) v = W[woff_c(l) + n * layer_kreal(l) + k];
    Wp[i] = (__bf16)v;
    const int ii = inf_index(l, n, k);
    if (ii >= 0) Wi[ii] = (__bf16)v;
  }
  if (i < BPAD_TOTAL) bp[i] = (i < B_TOTAL) ? b[i] : 0.f;
}

// ----------------------------------------------------------------------------
// Fused inference: all 5 layers of a block of 32 rows in one wave, weights
// staged in LDS once per workgroup.  Y [n][12] fp32.
//
// The product is formed transposed (weights as the A operand, activations
// as the B operand of v_mfma_f32_16x16x32_bf16), so a layer leaves lane
// (r, q) = (lane & 15, lane >> 4) holding features 16 b + 4 q + i (b = 0..7,
// i = 0..3) of row r of each 16-row block.  The next layer's B fragment for
// inputs 32 c .. 32 c + 31 wants, in lane (r, q), 8 inputs of row r: the
// hidden layers' weight columns are stored permuted (kperm) so that these 8
// are the lane's own features of blocks b = 2 c and 2 c + 1.  Activations
// therefore stay in registers from layer to layer (no LDS round trip, no
// barrier), and LDS holds only the weights: 121 KB, one workgroup per CU.
// A wave carries two 16-row blocks (each A fragment read from LDS feeds two
// MFMA) in 228 VGPRs, so the workgroup has 8 waves, two per SIMD: one
// wave's SiLU epilogue (two quarter-rate transcendentals per activation,
// the kernel's largest cost) overlaps the other's MFMA.  Waves take 32-row
// blocks round robin over the grid (block k on workgroup k mod grid), so a
// short list spreads over every CU.
// ----------------------------------------------------------------------------
constexpr int INF_WAVES = 8;
constexpr int INF_THREADS = INF_WAVES * 64;
constexpr int INF_RB = 2;                        // 16-row MFMA blocks per wave block
constexpr int INF_ROWS = 16 * INF_RB;            // rows per wave block
__device__ __forceinline__ bf16x8 cat4(bf16x4 a, bf16x4 b) {
  return bf16x8{a[0], a[1], a[2], a[3], b[0], b[1], b[2], b[3]};
}

// A layer's accumulators start at its bias (lane (r, q) holds features
// 16 b + 4 q + i), so the MFMA chain leaves Z = W x + b and the epilogue
// adds nothing
__device__ __forceinline__ void bias_init(f32x4 (&acc)[INF_RB][8], const float* bias, int q) {
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const f32x4 bc = *reinterpret_cast<const f32x4*>(bias + 16 * b + 4 * q);
#pragma unroll
    for (int a = 0; a < INF_RB; ++a) acc[a][b] = bc;
  }
}

// SiLU of one layer's Z into the next layer's inputs
__device__ __forceinline__ void hidden_epilogue(const f32x4 (&acc)[INF_RB][8], bf16x4 (&h)[INF_RB][8]) {
#pragma unroll
  for (int b = 0; b < 8; ++b)
#pragma unroll
    for (int a = 0; a < INF_RB; ++a)
#pragma unroll
      for (int i = 0; i < 4; ++i) h[a][b][i] = (__bf16)silu(acc[a][b][i]);
}

__global__ void __launch_bounds__(INF_THREADS) mlp_infer(const __bf16* __restrict__ Wi,
                                                         const float* __restrict__ bp,
                                                         const float* __restrict__ X, int64_t n_host,
                                                         const unsigned* __restrict__ seg_count,
                                                         int64_t seg_cap, float* __restrict__ Y) {
  // rows 0 .. n_host - 1 of X and Y, or the step's list: its segment
  // counters are read on the device (graph-captured steps), the i-th listed
  // point's row is its slot
  __shared__ unsigned off[LIST_SEGS + 1];
  const int64_t n = seg_count ? (int64_t)list_scan(seg_count, off) : n_host;
  const int64_t nblk = (n + INF_ROWS - 1) / INF_ROWS;
  // a workgroup without a row block leaves before copying the weights
  if ((int64_t)blockIdx.x >= nblk) return;
  extern __shared__ __align__(16) unsigned char smem[];
  __bf16* sW = reinterpret_cast<__bf16*>(smem);
  float* sb = reinterpret_cast<float*>(smem + WLDS_TOTAL * sizeof(__bf16));
  const int tid = threadIdx.x;
  // the inference layout as it is, 16 bytes per copy
  static_assert(WLDS_TOTAL % 8 == 0, "whole 16-byte chunks");
  for (int t = tid; t < WLDS_TOTAL / 8; t += blockDim.x)
    reinterpret_cast<bf16x8*>(sW)[t] = reinterpret_cast<const bf16x8*>(Wi)[t];
  for (int i = tid; i < BLDS_TOTAL; i += blockDim.x) sb[i] = bp[i];
  __syncthreads();
  const int lane = tid & 63, wave = tid >> 6;
  const int r = lane & 15, q = lane >> 4;
  const int64_t nwaves = (int64_t)gridDim.x * INF_WAVES;
  static_assert(INF_ROWS <= 64, "one lane holds the slot of each row of a block");
  for (int64_t blk = (int64_t)wave * gridDim.x + blockIdx.x; blk < nblk; blk += nwaves) {
    const int64_t row0 = blk * INF_ROWS;
    // lane l < INF_ROWS: the slot of row row0 + l
    int slot_l = 0;
    if (lane < INF_ROWS && row0 + lane < n)
      slot_l = seg_count ? (int)list_slot(off, (unsigned)(row0 + lane), seg_cap) : (int)(row0 + lane);
    // layer 0's B fragments: inputs 8 q .. 8 q + 7 of row 16 a + r (NIN = 16
    // real inputs; lanes q >= 2 hold the zero padding)
    bf16x8 x0[INF_RB];
#pragma unroll
    for (int a = 0; a < INF_RB; ++a) {
      const int64_t sl = __shfl(slot_l, 16 * a + r);
      float4 lo = make_float4(0.f, 0.f, 0.f, 0.f), hi = lo;
      if (q < 2 && row0 + 16 * a + r < n) {
        const float4* src = reinterpret_cast<const float4*>(X + sl * NIN + 8 * q);
        lo = src[0];
        hi = src[1];
      }
      x0[a] = bf16x8{(__bf16)lo.x, (__bf16)lo.y, (__bf16)lo.z, (__bf16)lo.w,
                     (__bf16)hi.x, (__bf16)hi.y, (__bf16)hi.z, (__bf16)hi.w};
    }
    bf16x4 h[INF_RB][8];
    f32x4 acc[INF_RB][8];
    // layer 0 (K = 32)
    bias_init(acc, sb + boff(0), q);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bf16x8 wf = *reinterpret_cast<const bf16x8*>(sW + (16 * b + r) * wld(0) + 8 * q);
#pragma unroll
      for (int a = 0; a < INF_RB; ++a)
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf, x0[a], acc[a][b], 0, 0, 0);
    }
    hidden_epilogue(acc, h);
    // hidden layers 1 .. NLAYER - 2 (K = 128)
#pragma unroll 1
    for (int l = 1; l < NLAYER - 1; ++l) {
      const __bf16* W = sW + woff_lds(1) + (l - 1) * NH * wld(1);
      bias_init(acc, sb + boff(l), q);
#pragma unroll
      for (int c = 0; c < NH / 32; ++c) {
        bf16x8 xf[INF_RB];
#pragma unroll
        for (int a = 0; a < INF_RB; ++a) xf[a] = cat4(h[a][2 * c], h[a][2 * c + 1]);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
          const bf16x8 wf =
              *reinterpret_cast<const bf16x8*>(W + (16 * b + r) * wld(1) + 32 * c + 8 * q);
#pragma unroll
          for (int a = 0; a < INF_RB; ++a)
            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf, xf[a], acc[a][b], 0, 0, 0);
        }
      }
      hidden_epilogue(acc, h);
    }
    // output layer: one 16-feature block, Y = Z (fp32)
    {
      constexpr int l = NLAYER - 1;
      const __bf16* W = sW + woff_lds(l);
      // features 4 q .. 4 q + 3 of the output block, from its bias (zero
      // beyond NOUT)
      const f32x4 bo = *reinterpret_cast<const f32x4*>(sb + boff(l) + 4 * q);
      f32x4 o[INF_RB];
#pragma unroll
      for (int a = 0; a < INF_RB; ++a) o[a] = bo;
#pragma unroll
      for (int c = 0; c < NH / 32; ++c) {
        const bf16x8 wf = *reinterpret_cast<const bf16x8*>(W + r * wld(l) + 32 * c + 8 * q);
#pragma unroll
        for (int a = 0; a < INF_RB; ++a)
          o[a] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf, cat4(h[a][2 * c], h[a][2 * c + 1]),
                                                         o[a], 0, 0, 0);
      }
#pragma unroll
      for (int a = 0; a < INF_RB; ++a) {
        const int64_t m = row0 + 16 * a + r;
        const int64_t sl = __shfl(slot_l, 16 * a + r);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int col = 4 * q + i;
          if (m < n && col < NOUT) Y[sl * NOUT + col] = o[a][i];
        }
      }
    }
  }
}

constexpr size_t INF_SMEM = WLDS_TOTAL * sizeof(__bf16) + BLDS_TOTAL * sizeof(float);

// splitmix64 counter-based generator
__device__ __forceinline__ uint64_t smix(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
struct Rng {
  uint64_t s;
  __device__ double uni() {  // (0, 1)
    s = smix(s);
    return ((double)(s >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  }
  __device__ double normal() {
    const double u1 = uni(), u2 = uni();
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
  }
};

// Principal values (descending) of a symmetric 3x3 log-strain tensor.
__device__ __forceinline__ void principal(const double (&s)[6], double& l0, double& l1, double& l2) {
  double q[9];
  sma::eig3<double>(s[0], s[3], s[4], s[1], s[5], s[2], l0, l1, l2, q);
}

struct GenArgs {
  int64_t n;
  uint64_t seed;
  const double* props;     // base material [18]
  double strain_sd, step_sd, jitter;
  float* X;                // [n][16]
  float* T;                // [n][12]
  uint32_t* flags;         // [n]
};

// Training samples drawn like the material-point workload: log-strain components ~ N(0, sd^2),
// the previous state one increment (~ N(0, step_sd^2)) away, xS_0 in [0, 1]
// (the transformation strain eL xS_0 inside the 0.05 ball), T ~ U[290, 340],
// material parameters jittered +-jitter.  Labels from the exact FP64 map.
__global__ void __launch_bounds__(256) gen_samples(GenArgs G) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= G.n) return;
  Rng rng{smix(G.seed * 0x2545F4914F6CDD1Dull + (uint64_t)i)};
  double p[18];
#pragma unroll
  for (int k = 0; k < 18; ++k) p[k] = G.props[k];
  const int jit[7] = {0, 1, 2, 3, 4, 9, 10};  // E, nu, eL, sigma_s/f AS, sigma_s/f SA
#pragma unroll
  for (int k = 0; k < 7; ++k) p[jit[k]] *= 1.0 + G.jitter * (2.0 * rng.uni() - 1.0);
  const double E = p[0];
  const double T = 290.0 + 50.0 * rng.uni();
  double eps[6], prev[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    eps[k] = G.strain_sd * rng.normal();
    prev[k] = eps[k] - G.step_sd * rng.normal();
  }
  const double u = rng.uni();
  double xS0 = (u < 0.15) ? 0.0 : (u < 0.25) ? 1.0 : rng.uni();
  double l0, l1, l2, q0, q1, q2;
  principal(eps, l0, l1, l2);
  principal(prev, q0, q1, q2);
  const double th = l0 + l1 + l2, thp = q0 + q1 + q2;
  const double e0 = l0 - th / 3.0, e1 = l1 - th / 3.0, e2 = l2 - th / 3.0;
  const double ep0 = q0 - thp / 3.0, ep1 = q1 - thp / 3.0, ep2 = q2 - thp / 3.0;
  const double en = sqrt(e0 * e0 + e1 * e1 + e2 * e2);
  const double enp = sqrt(ep0 * ep0 + ep1 * ep1 + ep2 * ep2);
  const sma::Mat m = sma::make_mat(p, E);
  const double FAS0 = sma::loading(m, m.C_AS, thp, enp, xS0);
  const double FSA0 = sma::loading(m, m.C_SA, thp, enp, xS0);
  const double Ftr = sma::loading(m, m.C_AS, th, en, xS0);
  const sma::ReturnMap rm = sma::return_map(m, th, en, xS0, FAS0, FSA0);
  float* f = G.X + i * NIN;
  make_features(p, E, T, e0, e1, e2, th, xS0, FAS0, FSA0, Ftr, f);
  const double d0 = en > 0 ? e0 / en : 0.0, d1 = en > 0 ? e1 / en : 0.0, d2 = en > 0 ? e2 / en : 0.0;
  const double vol = m.K * (th - 3.0 * m.alfa * m.eL * rm.xS);
  float* t = G.T + i * NOUT;
  t[0] = (float)((vol + 2.0 * m.G * (e0 - m.eL * rm.xS * d0)) * Scale::tau);
  t[1] = (float)((vol + 2.0 * m.G * (e1 - m.eL * rm.xS * d1)) * Scale::tau);
  t[2] = (float)((vol + 2.0 * m.G * (e2 - m.eL * rm.xS * d2)) * Scale::tau);
  t[3] = (float)(m.eL * rm.xS * d0 * Scale::etr);
  t[4] = (float)(m.eL * rm.xS * d1 * Scale::etr);
  t[5] = (float)(m.eL * rm.xS * d2 * Scale::etr);
  t[6] = (float)rm.xS;
  t[7] = (float)(rm.F_AS * Scale::F);
  t[8] = (float)(rm.F_SA * Scale::F);
  t[9] = (rm.flags & sma::FLAG_KF_AS) ? 1.f : 0.f;
  t[10] = (rm.flags & sma::FLAG_KF_SA) ? 1.f : 0.f;
  t[11] = (float)((rm.xS - xS0) * Scale::dx);
  G.flags[i] = rm.flags;
}

}  // namespace mlp
