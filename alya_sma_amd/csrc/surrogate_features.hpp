// Input features and output targets of the MLP surrogate (shared by the
// step kernels and the surrogate's own kernels).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mlp {

constexpr int NIN = 16, NH = 128, NOUT = 12;

// ----------------------------------------------------------------------------
// Features and targets.  Inputs (16): principal deviatoric log strains
// e_1..3 (descending), theta, xS_0, F_AS_0, F_SA_0, T, E, nu, eL, sigma_s/f
// for AS and SA, and the trial loading value F_AS(xS_0), each with a fixed
// affine scaling.  Targets (12): principal Kirchhoff stresses, principal
// transformation strains eL xS d_i, xS, F_AS, F_SA, the two branch flags and
// (xS - xS_0).
// ----------------------------------------------------------------------------
struct Scale {
  static constexpr float e = 1.f / 0.05f, th = 1.f / 0.02f, F = 1.f / 1000.f, tau = 1.f / 1000.f;
  static constexpr float etr = 1.f / 0.05f, dx = 1.f;
};

__host__ __device__ __forceinline__ void make_features(const double* p, double E, double T, double e0,
                                                       double e1, double e2, double theta,
                                                       double xS0, double FAS0, double FSA0,
                                                       double Ftrial, float* f) {
  f[0] = (float)(e0 * Scale::e);
  f[1] = (float)(e1 * Scale::e);
  f[2] = (float)(e2 * Scale::e);
  f[3] = (float)(theta * Scale::th);
  f[4] = (float)xS0;
  f[5] = (float)(FAS0 * Scale::F);
  f[6] = (float)(FSA0 * Scale::F);
  f[7] = (float)((T - 315.0) / 25.0);
  f[8] = (float)((E - 60000.0) / 6000.0);
  f[9] = (float)((p[1] - 0.33) / 0.033);
  f[10] = (float)((p[2] - 0.05) / 0.005);
  f[11] = (float)((p[3] - 400.0) / 40.0);
  f[12] = (float)((p[4] - 500.0) / 50.0);
  f[13] = (float)((p[9] - 250.0) / 25.0);
  f[14] = (float)((p[10] - 150.0) / 15.0);
  f[15] = (float)(Ftrial * Scale::F);
}

// ----------------------------------------------------------------------------
// The step's list of points for the network.  Appends go to LIST_SEGS
// segments, each with its own counter (counters LIST_STRIDE words, 128
// bytes, apart): wave w of the listing pass appends to segment
// w % LIST_SEGS, so the appends of the waves spread over LIST_SEGS
// addresses instead of queueing on one device-wide atomic.  Segment s owns
// slots [s * seg_cap, s * seg_cap + count_s); seg_cap = 64 x the number of
// waves that append to one segment (a wave lists at most its 64 points).
// The readers number the listed points densely, 0 .. total - 1, across the
// segments in order.
// ----------------------------------------------------------------------------
constexpr int LIST_SEGS = 64;
constexpr int LIST_STRIDE = 32;

__host__ __device__ constexpr int64_t list_seg_cap(int64_t nwaves) {
  return 64 * ((nwaves + LIST_SEGS - 1) / LIST_SEGS);
}

// Exclusive prefix sums of the segment counts into off[0 .. LIST_SEGS]
// (off[LIST_SEGS] = total), formed by the block's first wave; every thread
// of the block must call it (it ends with a barrier).
__device__ inline unsigned list_scan(const unsigned* count, unsigned* off) {
  static_assert(LIST_SEGS % 64 == 0, "whole lanes of segments");
  constexpr int PER = LIST_SEGS / 64;   // segments per lane
  if (threadIdx.x < 64) {
    const int l = threadIdx.x;
    unsigned c[PER], sum = 0;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      c[j] = count[(l * PER + j) * LIST_STRIDE];
      sum += c[j];
    }
    unsigned v = sum;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const unsigned u = __shfl_up(v, d);
      if (l >= d) v += u;
    }
    unsigned o = v - sum;
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      off[l * PER + j] = o;
      o += c[j];
    }
    if (l == 63) off[LIST_SEGS] = v;
  }
  __syncthreads();
  return off[LIST_SEGS];
}

// Slot of the i-th listed point (i < off[LIST_SEGS]): the segment s with
// off[s] <= i < off[s + 1] (binary search; empty segments are passed over).
__device__ inline int64_t list_slot(const unsigned* off, unsigned i, int64_t seg_cap) {
  int lo = 0, hi = LIST_SEGS;   // off[lo] <= i < off[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= i) lo = mid;
    else hi = mid;
  }
  return (int64_t)lo * seg_cap + (i - off[lo]);
}

}  // namespace mlp
