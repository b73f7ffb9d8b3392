// Input features and output targets of the MLP surrogate (shared by the
// step kernels and the surrogate's own kernels).
#pragma once
#include <hip/hip_runtime.h>

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

}  // namespace mlp
