// Device-side superelastic SMA law (Alya solidz stress model 160) for gfx950.
//
// Reference: sld_stress_model_160.f90 (compute_kirchoff_stress :212-500,
// eval_loading_function :513-537, eval_beta :539-556 with beta forced to 1,
// INV2x2 :558-573), dsyevq3.f90:22-137 + dsytrd3.f90:22-109 (Kopp's
// Householder + QL), eigsrt.f90:19-48 (descending selection sort).
//
// Two spectral front ends feed one FP64 return map:
//   * spectral_fp64: b = F F^T, Kopp QL on b, sqrt/pow/log exactly as the
//     reference sequences them (plain FP64 mode, parity <= 1e-10);
//   * spectral_fp32: the same QL on b - I = H + H^T + H H^T (H = F - I) in
//     FP32, with e_i = log1p(mu_i)/2 - theta/3.  Working on b - I keeps the
//     small strains at full FP32 relative precision (FP32 kinematics mode).
// The loading functions, branch flags and the 2x2 Newton on (dh_AS, dh_SA)
// always run in FP64, with FMA contraction off so that every branch decision
// sees the same rounding as the reference.
//
// All loops over the 3 eigenpairs are written with compile-time indices so
// that the state stays in VGPRs (no dynamically indexed private arrays).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sma {

constexpr uint32_t FLAG_KF_AS = 0x01u;
constexpr uint32_t FLAG_KF_SA = 0x02u;
constexpr uint32_t FLAG_NEWTON_FAIL = 0x04u;
constexpr uint32_t FLAG_EIGEN_FAIL = 0x08u;
constexpr uint32_t FLAG_DETF_NONPOS = 0x10u;
constexpr uint32_t FLAG_SURR_USED = 0x20u;
constexpr uint32_t FLAG_SURR_FALLBACK = 0x40u;
constexpr uint32_t FLAG_BAD_INPUT = 0x80u;
constexpr int FLAG_ITER_SHIFT = 8;

// Material constants derived once per point from props(18) and E
// (compute_kirchoff_stress :263-289).
struct Mat {
  double K, G, eL, alfa, Rf_AS, Rs_AS, Rf_SA, Rs_SA, G_1, C_AS, C_SA;
};

// Bulk and shear modulus from Young's modulus E and props(2) = nu.
__host__ __device__ __forceinline__ void elastic_moduli(const double* __restrict__ p, double E,
                                                       double& K, double& G) {
#pragma clang fp contract(off)
  const double v = p[1];
  K = E / (3.0 - 6.0 * v);
  G = E / (2.0 + 2.0 * v);
}

// Material constants with K, G given (precomputed per element).
__host__ __device__ __forceinline__ Mat make_mat_kg(const double* __restrict__ p, double K, double G) {
#pragma clang fp contract(off)
  Mat m;
  m.eL = p[2];
  m.alfa = p[14];
  m.C_AS = p[5];
  m.C_SA = p[11];
  m.K = K;
  m.G = G;
  double aux = sqrt(2.0 / 3.0) + m.alfa;
  m.Rf_AS = p[4] * aux + p[5] * p[6];
  m.Rs_AS = p[3] * aux + p[5] * p[7];
  m.Rf_SA = p[10] * aux + p[11] * p[12];
  m.Rs_SA = p[9] * aux + p[11] * p[13];
  m.G_1 = (2.0 * m.G + 9.0 * (m.alfa * m.alfa) * m.K) * m.eL;
  return m;
}

// Constants of make_mat_kg that depend on props(18) alone, per material:
// c[0..3] = Rf_AS, Rs_AS, Rf_SA, Rs_SA (the same operations).
__host__ __device__ __forceinline__ void mat_thresholds(const double* __restrict__ p, double* c) {
  const Mat m = make_mat_kg(p, 0.0, 0.0);
  c[0] = m.Rf_AS;
  c[1] = m.Rs_AS;
  c[2] = m.Rf_SA;
  c[3] = m.Rs_SA;
}

// make_mat_kg with the thresholds c = mat_thresholds(p) given.
__host__ __device__ __forceinline__ Mat make_mat_pre(const double* __restrict__ p,
                                                    const double* __restrict__ c, double K,
                                                    double G) {
#pragma clang fp contract(off)
  Mat m;
  m.eL = p[2];
  m.alfa = p[14];
  m.C_AS = p[5];
  m.C_SA = p[11];
  m.K = K;
  m.G = G;
  m.Rf_AS = c[0];
  m.Rs_AS = c[1];
  m.Rf_SA = c[2];
  m.Rs_SA = c[3];
  m.G_1 = (2.0 * m.G + 9.0 * (m.alfa * m.alfa) * m.K) * m.eL;
  return m;
}

__host__ __device__ __forceinline__ Mat make_mat(const double* __restrict__ p, double E) {
  double K, G;
  elastic_moduli(p, E, K, G);
  return make_mat_kg(p, K, G);
}

// 1/x: the hardware reciprocal (1 ulp) on the device, for FP32-only paths.
__host__ __device__ __forceinline__ float fast_rcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}

// J^(-2/3) in FP64 (the neo-Hookean law's volumetric factor).  On the
// device: y = J^(-1/3) from the hardware single-precision log2 and exp2
// (relative error ~3e-7), refined by two Newton steps on y^-3 = J,
// y <- y + y (1 - J y^3) / 3 (error e -> -2 e^2: 3e-7 -> 2e-13 -> rounding),
// then squared: within 2 ulps of the exact value (pow: 1), in ~20
// instructions against pow's ~170.  J outside [2^-100, 2^100], NaN
// included, takes pow.
__host__ __device__ __forceinline__ double pow_m23(double J) {
#if defined(__HIP_DEVICE_COMPILE__)
  if (!(J >= 0x1p-100 && J <= 0x1p100)) return pow(J, -2.0 / 3.0);
  const float l2 = __builtin_amdgcn_logf((float)J);
  double y = (double)__builtin_amdgcn_exp2f(l2 * (-1.0f / 3.0f));
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const double r = fma(-J * y, y * y, 1.0);
    y = fma(y * (1.0 / 3.0), r, y);
  }
  return y * y;
#else
  return pow(J, -2.0 / 3.0);
#endif
}

// One Givens step of the implicit-shift QL sweep (dsyevq3.f90:99-129) on
// scalar state: off-diagonals ej, ej1, diagonals wj, wj1, eigenvector columns
// (a0,a1,a2) = column j and (b0,b1,b2) = column j+1.  Scalars, not arrays, so
// that every value stays in a VGPR.
template <typename R>
__host__ __device__ __forceinline__ void ql_rot(R& s, R& c, R& p, R& g, R ej, R& ej1, R wj, R& wj1,
                                       R& a0, R& a1, R& a2, R& b0, R& b1, R& b2) {
#pragma clang fp contract(off)
  R f = s * ej;
  R b = c * ej;
  // Both branches of :102-114 as one select (identical arithmetic per branch):
  // |f| > |g|: c = g/f, r = sqrt(1+c^2), e = f r, s = 1/r, c = c s
  // else     : s = f/g, r = sqrt(1+s^2), e = g r, c = 1/r, s = s c
  const bool big = fabs(f) > fabs(g);
  const R num = big ? g : f, den = big ? f : g;
  const R t0 = num / den;
  R r = sqrt(R(1) + t0 * t0);
  ej1 = den * r;
  const R inv = R(1) / r;
  const R tr = t0 * inv;
  c = big ? tr : inv;
  s = big ? inv : tr;
  g = wj1 - p;
  r = (wj - g) * s + R(2) * c * b;
  p = s * r;
  wj1 = g + p;
  g = c * r - b;
  R t;
  t = b0; b0 = s * a0 + c * t; a0 = c * a0 - s * t;
  t = b1; b1 = s * a1 + c * t; a1 = c * a1 - s * t;
  t = b2; b2 = s * a2 + c * t; a2 = c * a2 - s * t;
}

// Shift of dsyevq3.f90:87-94 for (L, M): wl, wl1 = W(L), W(L+1); wm = W(M).
template <typename R>
__host__ __device__ __forceinline__ R ql_shift(R wl, R wl1, R wm, R el) {
#pragma clang fp contract(off)
  R g = (wl1 - wl) / (R(2) * el);
  R r = sqrt(R(1) + g * g);
  if (g >= R(0)) return wm - wl + el / (g + r);
  return wm - wl + el / (g - r);
}

// Symmetric 3x3 eigen-decomposition: dsytrd3 + dsyevq3 + eigsrt (largest
// eigenvalue first).  Eigenvectors in q[9] column-major.  Returns false on
// the 30-iteration bail-out (dsyevq3.f90:82-84).
template <typename R>
__host__ __device__ __forceinline__ bool eig3(R a00, R a01, R a02, R a11, R a12, R a22,
                                     R& w0, R& w1, R& w2, R (&q)[9]) {
#pragma clang fp contract(off)
  // dsytrd3.f90:22-109 (Q starts as the identity)
  R q00 = 1, q01 = 0, q02 = 0, q10 = 0, q11 = 1, q12 = 0, q20 = 0, q21 = 0, q22 = 1;
  R e0, e1, e2 = 0;
  R h = a01 * a01 + a02 * a02;
  R g = (a01 > R(0)) ? -sqrt(h) : sqrt(h);
  e0 = g;
  R f = g * a01;
  R u1 = a01 - g, u2 = a02;
  R omega = h - f;
  if (omega > R(0)) {
    omega = R(1) / omega;
    R f1 = a11 * u1 + a12 * u2;          // I = 2: A(2,2)U(2) + A(2,3)U(3)
    R p1 = omega * f1;
    R K = R(0) + u1 * f1;
    R f2 = a12 * u1 + a22 * u2;          // I = 3: A(2,3)U(2) + A(3,3)U(3)
    R p2 = omega * f2;
    K = K + u2 * f2;
    if constexpr (sizeof(R) == 8) K = R(0.5) * K * (omega * omega);   // reference order
    else K = R(0.5) * (K * omega) * omega;                            // no FP32 overflow
    p1 = p1 - K * u1;
    p2 = p2 - K * u2;
    w0 = a00;
    w1 = a11 - R(2) * p1 * u1;
    w2 = a22 - R(2) * p2 * u2;
    R fa = omega * u1;
    q11 = q11 - fa * u1;
    q21 = q21 - fa * u2;
    R fb = omega * u2;
    q12 = q12 - fb * u1;
    q22 = q22 - fb * u2;
    e1 = a12 - p1 * u2 - u1 * p2;
  } else {
    w0 = a00; w1 = a11; w2 = a22;
    e1 = a12;
  }
  // dsyevq3.f90:64-134: L = 1 (0-based 0), then L = 2
  bool ok = true;
  {
    int niter = 0;
    for (int it = 0; it < 50; ++it) {
      R g0 = fabs(w0) + fabs(w1);
      if (fabs(e0) + g0 == g0) break;                 // M = L
      R g1 = fabs(w1) + fabs(w2);
      const bool m1 = (fabs(e1) + g1 == g1);          // M = 2, else M = 3
      if (++niter >= 30) { ok = false; break; }
      R s = 1, c = 1, p = 0;
      if (m1) {
        R gg = ql_shift<R>(w0, w1, w1, e0);
        ql_rot<R>(s, c, p, gg, e0, e1, w0, w1, q00, q10, q20, q01, q11, q21);
        w0 = w0 - p; e0 = gg; e1 = 0;
      } else {
        R gg = ql_shift<R>(w0, w1, w2, e0);
        ql_rot<R>(s, c, p, gg, e1, e2, w1, w2, q01, q11, q21, q02, q12, q22);
        ql_rot<R>(s, c, p, gg, e0, e1, w0, w1, q00, q10, q20, q01, q11, q21);
        w0 = w0 - p; e0 = gg; e2 = 0;
      }
    }
  }
  if (ok) {
    int niter = 0;
    for (int it = 0; it < 50; ++it) {
      R g1 = fabs(w1) + fabs(w2);
      if (fabs(e1) + g1 == g1) break;
      if (++niter >= 30) { ok = false; break; }
      R s = 1, c = 1, p = 0;
      R gg = ql_shift<R>(w1, w2, w2, e1);
      ql_rot<R>(s, c, p, gg, e1, e2, w1, w2, q01, q11, q21, q02, q12, q22);
      w1 = w1 - p; e1 = gg; e2 = 0;
    }
  }
  // eigsrt.f90:19-48 as selects.  Pass 1 moves the last maximum ('>=') to
  // slot 1; pass 2 swaps slots 2 and 3 whenever d(3) >= d(2).
  {
    const bool k1 = (w1 >= w0) && !(w2 >= w1);
    const bool k2 = (w2 >= ((w1 >= w0) ? w1 : w0));
    const R pmax = k2 ? w2 : (k1 ? w1 : w0);
    const R nw1 = k1 ? w0 : w1, nw2 = k2 ? w0 : w2;
    R c00 = k1 ? q01 : (k2 ? q02 : q00), c10 = k1 ? q11 : (k2 ? q12 : q10),
      c20 = k1 ? q21 : (k2 ? q22 : q20);
    R c01 = k1 ? q00 : q01, c11 = k1 ? q10 : q11, c21 = k1 ? q20 : q21;
    R c02 = k2 ? q00 : q02, c12 = k2 ? q10 : q12, c22 = k2 ? q20 : q22;
    w0 = pmax;
    const bool sw = nw2 >= nw1;
    w1 = sw ? nw2 : nw1;
    w2 = sw ? nw1 : nw2;
    q00 = c00; q10 = c10; q20 = c20;
    q01 = sw ? c02 : c01; q11 = sw ? c12 : c11; q21 = sw ? c22 : c21;
    q02 = sw ? c01 : c02; q12 = sw ? c11 : c12; q22 = sw ? c21 : c22;
  }
  q[0] = q00; q[1] = q10; q[2] = q20;
  q[3] = q01; q[4] = q11; q[5] = q21;
  q[6] = q02; q[7] = q12; q[8] = q22;
  return ok;
}

// Log-strain invariants that the return map consumes, plus the principal
// directions and deviatoric strains the stress needs.
template <typename R>
struct Spectral {
  double theta, e_n;            // FP64 for the return map
  R e[3], d[3];                 // principal log strains and e_i / e_n
  R n[9];                       // principal directions (columns)
  bool eig_ok;
};

// Plain FP64: compute_kirchoff_stress :293-332 in reference order.
__host__ __device__ __forceinline__ Spectral<double> spectral_fp64(const double (&F)[9]) {
#pragma clang fp contract(off)
  Spectral<double> s;
  // b = matmul(F, transpose(F)) (col-major F: F(i,k) = F[i + 3k])
  double b00 = F[0] * F[0] + F[3] * F[3] + F[6] * F[6];
  double b01 = F[0] * F[1] + F[3] * F[4] + F[6] * F[7];
  double b02 = F[0] * F[2] + F[3] * F[5] + F[6] * F[8];
  double b11 = F[1] * F[1] + F[4] * F[4] + F[7] * F[7];
  double b12 = F[1] * F[2] + F[4] * F[5] + F[7] * F[8];
  double b22 = F[2] * F[2] + F[5] * F[5] + F[8] * F[8];
  double w0, w1, w2;
  s.eig_ok = eig3<double>(b00, b01, b02, b11, b12, b22, w0, w1, w2, s.n);
  double h1 = sqrt(w0), h2 = sqrt(w1), h3 = sqrt(w2);
  double J = h1 * h2 * h3;
  s.theta = log(J);
  double Jd = pow(J, -1.0 / 3.0);
  s.e[0] = log(Jd * h1);
  s.e[1] = log(Jd * h2);
  s.e[2] = log(Jd * h3);
  s.e_n = sqrt(s.e[0] * s.e[0] + s.e[1] * s.e[1] + s.e[2] * s.e[2]);
  const double eps = 2.220446049250313080847e-16;  // epsilon(0.0_rp)
  if (s.e_n > eps) {
    s.d[0] = s.e[0] / s.e_n; s.d[1] = s.e[1] / s.e_n; s.d[2] = s.e[2] / s.e_n;
  } else {
    s.d[0] = s.d[1] = s.d[2] = 0.0;
  }
  return s;
}

// FP32 kinematics: spectral decomposition of b - I from H = F - I.
__host__ __device__ __forceinline__ Spectral<float> spectral_fp32(const float (&H)[9]) {
  Spectral<float> s;
  // (b - I)_ij = H_ij + H_ji + sum_k H_ik H_jk
  float c00 = 2.f * H[0] + (H[0] * H[0] + H[3] * H[3] + H[6] * H[6]);
  float c11 = 2.f * H[4] + (H[1] * H[1] + H[4] * H[4] + H[7] * H[7]);
  float c22 = 2.f * H[8] + (H[2] * H[2] + H[5] * H[5] + H[8] * H[8]);
  float c01 = H[3] + H[1] + (H[0] * H[1] + H[3] * H[4] + H[6] * H[7]);
  float c02 = H[6] + H[2] + (H[0] * H[2] + H[3] * H[5] + H[6] * H[8]);
  float c12 = H[7] + H[5] + (H[1] * H[2] + H[4] * H[5] + H[7] * H[8]);
  // scale to unit magnitude so no intermediate of the FP32 Householder/QL
  // steps under- or overflows for strains from 1e-12 to 1
  float sc = fmaxf(fmaxf(fmaxf(fabsf(c00), fabsf(c11)), fmaxf(fabsf(c22), fabsf(c01))),
                   fmaxf(fabsf(c02), fabsf(c12)));
  // below 1e-30 the strain is zero to FP32 precision; a denormal scale would
  // overflow its reciprocal
  const float inv = sc > 1e-30f ? 1.0f / sc : 0.f;
  float mu0, mu1, mu2;
  s.eig_ok = eig3<float>(c00 * inv, c01 * inv, c02 * inv, c11 * inv, c12 * inv, c22 * inv,
                         mu0, mu1, mu2, s.n);
  mu0 *= sc;
  mu1 *= sc;
  mu2 *= sc;
  float l0 = 0.5f * log1pf(mu0), l1 = 0.5f * log1pf(mu1), l2 = 0.5f * log1pf(mu2);
  float th = l0 + l1 + l2;
  float t3 = th * (1.0f / 3.0f);
  s.e[0] = l0 - t3; s.e[1] = l1 - t3; s.e[2] = l2 - t3;
  float en = sqrtf(s.e[0] * s.e[0] + s.e[1] * s.e[1] + s.e[2] * s.e[2]);
  s.theta = (double)th;
  s.e_n = (double)en;
  if (en > 1.1920929e-07f) {
    float r = 1.0f / en;
    s.d[0] = s.e[0] * r; s.d[1] = s.e[1] * r; s.d[2] = s.e[2] * r;
  } else {
    s.d[0] = s.d[1] = s.d[2] = 0.f;
  }
  return s;
}

// FP32 kinematics without eigenvectors.  The stress of the law needs only
// theta = tr E, the deviatoric Hencky strain e' = dev E and |e'|, because
// tau = c_vol I + sum_i 2G (e_i - eL xS e_i/|e'|) n_i n_i^T
//     = c_vol I + 2G (1 - eL xS/|e'|) e'          (compute_kirchoff_stress
// :293-332 with d_i = e_i/|e'|).  E = 1/2 log b is formed as a tensor:
// 1/2 log b = artanh(Z) = sum_k Z^(2k+1)/(2k+1), Z = (b - I)(b + I)^-1, whose
// terms all commute with b; with |Z|_F <= 0.6 the terms fall by >= 0.36 each.
struct Hencky32 {
  double theta, e_n;   // FP64 for the return map
  float ed[6];         // e' (00, 01, 02, 11, 12, 22)
  bool ok;             // false: |Z|_F > 0.6, use the spectral path
};

// Principal values, descending, of a symmetric deviatoric 3x3 tensor
// s = (00, 01, 02, 11, 12, 22), in closed form: the roots of the
// characteristic polynomial l^3 - J2 l - J3 = 0 by the trigonometric
// solution, l_k = 2 r cos(a - 2 pi k / 3) with r = sqrt(J2 / 3),
// cos 3a = J3 / (2 r^3), a in [0, pi / 3] (k = 0, 1, 2 give the descending
// order).  Any trace left by rounding is removed first.  Below J2 = 1e-24
// (|s| ~ 1e-12) the values are returned as zero.
__host__ __device__ __forceinline__ void principal_dev_fp32(const float (&s)[6], float& l0, float& l1,
                                                           float& l2) {
  const float m = (s[0] + s[3] + s[5]) * (1.0f / 3.0f);
  const float d0 = s[0] - m, d1 = s[3] - m, d2 = s[5] - m;
  const float J2 = 0.5f * (d0 * d0 + d1 * d1 + d2 * d2) + s[1] * s[1] + s[2] * s[2] + s[4] * s[4];
  if (!(J2 > 1e-24f)) {
    l0 = l1 = l2 = 0.f;
    return;
  }
  const float J3 = d0 * (d1 * d2 - s[4] * s[4]) - s[1] * (s[1] * d2 - s[4] * s[2]) +
                   s[2] * (s[1] * s[4] - d1 * s[2]);
  const float r = sqrtf(J2 * (1.0f / 3.0f));
  const float c3 = fminf(1.0f, fmaxf(-1.0f, J3 / (2.0f * r * r * r)));
  const float a = acosf(c3) * (1.0f / 3.0f);
  const float c = cosf(a), sn = sinf(a);
  const float h = 1.7320508075688772f;   // 2 sin(2 pi / 3)
  l0 = 2.0f * r * c;
  l1 = r * (h * sn - c);
  l2 = -r * (c + h * sn);
}

// c = a b for symmetric a, b that commute (the product is symmetric).
__host__ __device__ __forceinline__ void sym_mul_commuting(const float (&a)[6], const float (&b)[6],
                                                          float (&c)[6]) {
  c[0] = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
  c[1] = a[0] * b[1] + a[1] * b[3] + a[2] * b[4];
  c[2] = a[0] * b[2] + a[1] * b[4] + a[2] * b[5];
  c[3] = a[1] * b[1] + a[3] * b[3] + a[4] * b[4];
  c[4] = a[1] * b[2] + a[3] * b[4] + a[4] * b[5];
  c[5] = a[2] * b[2] + a[4] * b[4] + a[5] * b[5];
}

__host__ __device__ __forceinline__ Hencky32 hencky_fp32(const float (&H)[9]) {
  Hencky32 h;
  // C = b - I from H = F - I without cancellation
  float c[6];
  c[0] = 2.f * H[0] + (H[0] * H[0] + H[3] * H[3] + H[6] * H[6]);
  c[3] = 2.f * H[4] + (H[1] * H[1] + H[4] * H[4] + H[7] * H[7]);
  c[5] = 2.f * H[8] + (H[2] * H[2] + H[5] * H[5] + H[8] * H[8]);
  c[1] = H[3] + H[1] + (H[0] * H[1] + H[3] * H[4] + H[6] * H[7]);
  c[2] = H[6] + H[2] + (H[0] * H[2] + H[3] * H[5] + H[6] * H[8]);
  c[4] = H[7] + H[5] + (H[1] * H[2] + H[4] * H[5] + H[7] * H[8]);
  // (b + I)^-1 = (2I + C)^-1 by cofactors (det(b + I) >= 1)
  const float m00 = 2.f + c[0], m11 = 2.f + c[3], m22 = 2.f + c[5];
  float a[6];
  a[0] = m11 * m22 - c[4] * c[4];
  a[1] = c[2] * c[4] - c[1] * m22;
  a[2] = c[1] * c[4] - c[2] * m11;
  a[3] = m00 * m22 - c[2] * c[2];
  a[4] = c[1] * c[2] - m00 * c[4];
  a[5] = m00 * m11 - c[1] * c[1];
  const float r = fast_rcp(m00 * a[0] + c[1] * a[1] + c[2] * a[2]);
#pragma unroll
  for (int q = 0; q < 6; ++q) a[q] *= r;
  float z[6];
  sym_mul_commuting(c, a, z);
  const float zf2 = z[0] * z[0] + z[3] * z[3] + z[5] * z[5] +
                    2.f * (z[1] * z[1] + z[2] * z[2] + z[4] * z[4]);
  h.ok = zf2 <= 0.36f;
  // terms after Z^(2n+1) sum to below 2^-24 of |E| ~ |Z| when
  // |Z|^(2n+2) / ((2n+3)(1 - |Z|^2)) <= 2^-24 (|Z| <= |Z|_F)
  const int nterms = zf2 <= 1.7e-7f ? 0 : zf2 <= 5.4e-4f ? 1 : zf2 <= 7.3e-3f ? 2
                   : zf2 <= 2.6e-2f ? 3 : zf2 <= 5.2e-2f ? 4 : zf2 <= 8.4e-2f ? 5
                   : zf2 <= 0.115f ? 6 : zf2 <= 0.17f ? 8 : zf2 <= 0.22f ? 10
                   : zf2 <= 0.28f ? 14 : 20;
  float z2[6], t[6], e[6];
  sym_mul_commuting(z, z, z2);
#pragma unroll
  for (int q = 0; q < 6; ++q) t[q] = e[q] = z[q];
  if (h.ok) {
    // written out term by term: constant weights 1/(2k+1), no loop-carried
    // register copies; each lane takes its own number of terms
    auto term = [&](float w) {
      float tn[6];
      sym_mul_commuting(t, z2, tn);
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        t[q] = tn[q];
        e[q] += w * tn[q];
      }
    };
    if (nterms >= 1) { term(1.0f / 3.0f);
    if (nterms >= 2) { term(1.0f / 5.0f);
    if (nterms >= 3) { term(1.0f / 7.0f);
    if (nterms >= 4) { term(1.0f / 9.0f);
    if (nterms >= 5) { term(1.0f / 11.0f);
    if (nterms >= 6) { term(1.0f / 13.0f);
    if (nterms >= 7) {
      for (int k = 7; k <= nterms; ++k) term(1.0f / (float)(2 * k + 1));
    } } } } } } }
  }
  const float th = e[0] + e[3] + e[5];
  const float t3 = th * (1.0f / 3.0f);
  h.ed[0] = e[0] - t3;
  h.ed[1] = e[1];
  h.ed[2] = e[2];
  h.ed[3] = e[3] - t3;
  h.ed[4] = e[4];
  h.ed[5] = e[5] - t3;
  const float en = sqrtf(h.ed[0] * h.ed[0] + h.ed[3] * h.ed[3] + h.ed[5] * h.ed[5] +
                         2.f * (h.ed[1] * h.ed[1] + h.ed[2] * h.ed[2] + h.ed[4] * h.ed[4]));
  h.theta = (double)th;
  h.e_n = (double)en;
  return h;
}

// FP64 Hencky strain by the same series, for the plain FP64 element step:
// E = 1/2 log b = sum_k Z^(2k+1)/(2k+1), Z = (b - I)(b + I)^-1, b - I formed
// from H = F - I.  theta = tr E and e' = dev E are what the loading functions
// and the stress consume (theta = log J, |e'| = e_n of compute_kirchoff_stress
// :313-322), so no eigenvectors, square roots of eigenvalues, logarithms or
// powers are needed; the terms run until the remainder is below 2^-54 of
// |Z| (per lane).  Stretches with |Z|_F > 0.6 (principal stretches beyond
// about 1/2..2) first halve the logarithm by symmetric square roots,
// log b = 2^k log b^(1/2^k), until |Z|_F <= 0.6.
struct Hencky64 {
  double theta, e_n;
  double ed[6];        // e' (00, 01, 02, 11, 12, 22)
  bool ok;             // false: no finite strain (a non-finite or non-positive b)
};

__host__ __device__ __forceinline__ void sym_mul_commuting64(const double (&a)[6],
                                                            const double (&b)[6], double (&c)[6]) {
  c[0] = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
  c[1] = a[0] * b[1] + a[1] * b[3] + a[2] * b[4];
  c[2] = a[0] * b[2] + a[1] * b[4] + a[2] * b[5];
  c[3] = a[1] * b[1] + a[3] * b[3] + a[4] * b[4];
  c[4] = a[1] * b[2] + a[3] * b[4] + a[4] * b[5];
  c[5] = a[2] * b[2] + a[4] * b[4] + a[5] * b[5];
}

// Inverse of a symmetric 3x3 matrix (00, 01, 02, 11, 12, 22) by its adjugate.
__host__ __device__ __forceinline__ void sym_inv64(const double (&m)[6], double (&r)[6]) {
  r[0] = m[3] * m[5] - m[4] * m[4];
  r[1] = m[2] * m[4] - m[1] * m[5];
  r[2] = m[1] * m[4] - m[2] * m[3];
  r[3] = m[0] * m[5] - m[2] * m[2];
  r[4] = m[1] * m[2] - m[0] * m[4];
  r[5] = m[0] * m[3] - m[1] * m[1];
  const double d = 1.0 / (m[0] * r[0] + m[1] * r[1] + m[2] * r[2]);
#pragma unroll
  for (int q = 0; q < 6; ++q) r[q] *= d;
}

// Z = (b - I)(b + I)^-1 from c = b - I; returns |Z|_F^2.
__host__ __device__ __forceinline__ double cayley64(const double (&c)[6], double (&z)[6]) {
  const double m[6] = {2.0 + c[0], c[1], c[2], 2.0 + c[3], c[4], 2.0 + c[5]};
  double a[6];
  sym_inv64(m, a);
  sym_mul_commuting64(c, a, z);
  return z[0] * z[0] + z[3] * z[3] + z[5] * z[5] + 2.0 * (z[1] * z[1] + z[2] * z[2] + z[4] * z[4]);
}

// Square root of a symmetric positive definite 3x3 matrix by the
// Denman-Beavers iteration Y <- (Y + W^-1)/2, W <- (W + Y^-1)/2 (Y -> b^1/2,
// W -> b^-1/2), run to convergence (quadratic; ~6-10 iterations for the
// stretches a solid reaches).
__host__ __device__ __forceinline__ void sym_sqrt64(double (&y)[6]) {
  double w[6] = {1.0, 0.0, 0.0, 1.0, 0.0, 1.0};
  for (int it = 0; it < 40; ++it) {
    double yi[6], wi[6];
    sym_inv64(y, yi);
    sym_inv64(w, wi);
    double d2 = 0.0, n2 = 0.0;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const double yn = 0.5 * (y[q] + wi[q]);
      const double f = (q == 1 || q == 2 || q == 4) ? 2.0 : 1.0;
      d2 += f * (yn - y[q]) * (yn - y[q]);
      n2 += f * yn * yn;
      y[q] = yn;
      w[q] = 0.5 * (w[q] + yi[q]);
    }
    if (!(d2 > 1e-32 * n2)) break;   // converged (or not finite)
  }
}

// 1/(2k+1), k = 0..40, the series' coefficients (rounded once, here, instead
// of a division per term on the device)
constexpr double kInvOdd64[41] = {
    1.0 / 1,  1.0 / 3,  1.0 / 5,  1.0 / 7,  1.0 / 9,  1.0 / 11, 1.0 / 13, 1.0 / 15, 1.0 / 17,
    1.0 / 19, 1.0 / 21, 1.0 / 23, 1.0 / 25, 1.0 / 27, 1.0 / 29, 1.0 / 31, 1.0 / 33, 1.0 / 35,
    1.0 / 37, 1.0 / 39, 1.0 / 41, 1.0 / 43, 1.0 / 45, 1.0 / 47, 1.0 / 49, 1.0 / 51, 1.0 / 53,
    1.0 / 55, 1.0 / 57, 1.0 / 59, 1.0 / 61, 1.0 / 63, 1.0 / 65, 1.0 / 67, 1.0 / 69, 1.0 / 71,
    1.0 / 73, 1.0 / 75, 1.0 / 77, 1.0 / 79, 1.0 / 81};

__host__ __device__ __forceinline__ Hencky64 hencky_fp64(const double (&H)[9]) {
  Hencky64 h;
  double c[6];
  c[0] = 2.0 * H[0] + (H[0] * H[0] + H[3] * H[3] + H[6] * H[6]);
  c[3] = 2.0 * H[4] + (H[1] * H[1] + H[4] * H[4] + H[7] * H[7]);
  c[5] = 2.0 * H[8] + (H[2] * H[2] + H[5] * H[5] + H[8] * H[8]);
  c[1] = H[3] + H[1] + (H[0] * H[1] + H[3] * H[4] + H[6] * H[7]);
  c[2] = H[6] + H[2] + (H[0] * H[2] + H[3] * H[5] + H[6] * H[8]);
  c[4] = H[7] + H[5] + (H[1] * H[2] + H[4] * H[5] + H[7] * H[8]);
  double z[6];
  double zf2 = cayley64(c, z);
  // large stretches: b <- b^(1/2) until |Z|_F <= 0.6 (each halves log b)
  int halvings = 0;
  if (zf2 > 0.36) {
    double b[6] = {1.0 + c[0], c[1], c[2], 1.0 + c[3], c[4], 1.0 + c[5]};
    while (zf2 > 0.36 && halvings < 10) {
      sym_sqrt64(b);
      ++halvings;
      const double cb[6] = {b[0] - 1.0, b[1], b[2], b[3] - 1.0, b[4], b[5] - 1.0};
      zf2 = cayley64(cb, z);
    }
  }
  h.ok = zf2 <= 0.36;
  double e[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) e[q] = z[q];
  if (h.ok && zf2 > 0.0) {
    // n terms after Z: the remainder |Z|^(2n+3)/((2n+3)(1 - |Z|^2)) stays
    // below 2^-54 |Z| (|Z| <= |Z|_F)
    const double lim = 5.551115123125783e-17 * (1.0 - zf2);
    double z2[6], t[6];
    sym_mul_commuting64(z, z, z2);
#pragma unroll
    for (int q = 0; q < 6; ++q) t[q] = z[q];
    double pw = zf2;
    for (int k = 1; k <= 40 && pw > lim * (double)(2 * k + 1); ++k) {
      double tn[6];
      sym_mul_commuting64(t, z2, tn);
      const double w = kInvOdd64[k];
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        t[q] = tn[q];
        e[q] += w * tn[q];
      }
      pw *= zf2;
    }
  }
  if (halvings) {
    const double sc = (double)(1 << halvings);
#pragma unroll
    for (int q = 0; q < 6; ++q) e[q] *= sc;
  }
  const double th = e[0] + e[3] + e[5];
  const double t3 = th * (1.0 / 3.0);
  h.ed[0] = e[0] - t3;
  h.ed[1] = e[1];
  h.ed[2] = e[2];
  h.ed[3] = e[3] - t3;
  h.ed[4] = e[4];
  h.ed[5] = e[5] - t3;
  h.theta = th;
  h.e_n = sqrt(h.ed[0] * h.ed[0] + h.ed[3] * h.ed[3] + h.ed[5] * h.ed[5] +
               2.0 * (h.ed[1] * h.ed[1] + h.ed[2] * h.ed[2] + h.ed[4] * h.ed[4]));
  return h;
}

// The law's Kirchhoff stress from the Hencky tensor (FP64):
// tau = c_vol I + 2G (1 - eL xS/e_n) e', the direction term dropped for
// e_n <= epsilon(0.0_rp) as d_i = 0 in compute_kirchoff_stress :323-332.
__host__ __device__ __forceinline__ void kirchhoff_hencky64(const Mat& m, const Hencky64& h,
                                                           double xS, double (&t)[6]) {
  const double cvol = m.K * (h.theta - 3.0 * m.alfa * m.eL * xS);
  const double twoG = 2.0 * m.G;
  const double coef = h.e_n > 2.220446049250313080847e-16 ? twoG * (1.0 - m.eL * xS / h.e_n) : twoG;
  t[0] = cvol + coef * h.ed[0];
  t[1] = coef * h.ed[1];
  t[2] = coef * h.ed[2];
  t[3] = cvol + coef * h.ed[3];
  t[4] = coef * h.ed[4];
  t[5] = cvol + coef * h.ed[5];
}

// tau = c_vol I + 2G (1 - eL xS/|e'|) e'; for |e'| below FP32 epsilon the
// direction term is dropped, as d_i = 0 in the spectral form.
__host__ __device__ __forceinline__ void kirchhoff_dev(const Mat& m, const Hencky32& h, double xS,
                                                      float (&t)[6]) {
  const float cvol = (float)(m.K * (h.theta - 3.0 * m.alfa * m.eL * xS));
  const float twoG = (float)(2.0 * m.G);
  const float en = (float)h.e_n;
  const float coef = en > 1.1920929e-07f ? twoG * (1.0f - (float)(m.eL * xS) * fast_rcp(en)) : twoG;
  t[0] = cvol + coef * h.ed[0];
  t[1] = coef * h.ed[1];
  t[2] = coef * h.ed[2];
  t[3] = cvol + coef * h.ed[3];
  t[4] = coef * h.ed[4];
  t[5] = cvol + coef * h.ed[5];
}

// Neo-Hookean (Ansys form) Kirchhoff stress, sld_stress_model_102.f90:116-119
// pushed forward: tau = F S F^T with S = ((-mu/3) J^(-2/3) tr C + J K (J-1))
// C^-1 + mu J^(-2/3) I gives tau = mu J^(-2/3) dev(b) + J K (J-1) I.  Formed
// from H = F - I: dev(b) = dev(b - I) and J - 1 = tr H + I2(H) + det H, so
// small strains lose nothing to cancellation (in FP32 as in FP64).
template <typename R>
__host__ __device__ __forceinline__ void neo_hookean_tau(double mu, double K, const R (&H)[9],
                                                        R (&t)[6], R& jm1_out) {
  R c[6];   // b - I = H + H^T + H H^T
  c[0] = R(2) * H[0] + (H[0] * H[0] + H[3] * H[3] + H[6] * H[6]);
  c[3] = R(2) * H[4] + (H[1] * H[1] + H[4] * H[4] + H[7] * H[7]);
  c[5] = R(2) * H[8] + (H[2] * H[2] + H[5] * H[5] + H[8] * H[8]);
  c[1] = H[3] + H[1] + (H[0] * H[1] + H[3] * H[4] + H[6] * H[7]);
  c[2] = H[6] + H[2] + (H[0] * H[2] + H[3] * H[5] + H[6] * H[8]);
  c[4] = H[7] + H[5] + (H[1] * H[2] + H[4] * H[5] + H[7] * H[8]);
  // det(I + H) - 1 (H column-major: H(i,k) = H[i + 3k])
  const R trH = H[0] + H[4] + H[8];
  const R i2 = (H[0] * H[4] - H[3] * H[1]) + (H[0] * H[8] - H[6] * H[2]) + (H[4] * H[8] - H[7] * H[5]);
  const R dH = H[0] * (H[4] * H[8] - H[7] * H[5]) - H[3] * (H[1] * H[8] - H[7] * H[2]) +
               H[6] * (H[1] * H[5] - H[4] * H[2]);
  const R jm1 = trH + i2 + dH;
  const R J = R(1) + jm1;
  R jm23;
  if constexpr (sizeof(R) == 8) jm23 = pow_m23(J);
  else jm23 = expf((-2.0f / 3.0f) * log1pf(jm1));
  const R tr3 = (c[0] + c[3] + c[5]) * (R(1) / R(3));
  const R m = (R)mu * jm23;
  const R p = J * (R)K * jm1;
  t[0] = m * (c[0] - tr3) + p;
  t[3] = m * (c[3] - tr3) + p;
  t[5] = m * (c[5] - tr3) + p;
  t[1] = m * c[1];
  t[2] = m * c[2];
  t[4] = m * c[4];
  jm1_out = jm1;
}

// eval_loading_function :513-537 (beta = 1 per :554, temperature inert).
__host__ __device__ __forceinline__ double loading(const Mat& m, double C, double theta,
                                          double enor, double xS) {
#pragma clang fp contract(off)
  double pres = m.K * (theta - 3.0 * m.alfa * m.eL * xS);
  double trac = 2.0 * m.G * (enor - m.eL * xS);
  return trac * 1.0 + 3.0 * m.alfa * pres - C * 0.0;
}

// Branch flags and return map (:334-484).  sdv in = TIME_N (xS_0, F_AS_0,
// F_SA_0); sdv out = ITER_K.  Returns status bits; xS is the clamped value.
struct ReturnMap {
  double xS, F_AS, F_SA;
  uint32_t flags;
};

__host__ __device__ __forceinline__ bool flags_only(const Mat& m, double theta, double e_n,
                                           double xS_0, double F_AS_0, double F_SA_0,
                                           bool& kf_AS, bool& kf_SA,
                                           double& F_AS, double& F_SA) {
#pragma clang fp contract(off)
  F_AS = loading(m, m.C_AS, theta, e_n, xS_0);
  kf_AS = (F_AS > F_AS_0) && (F_AS - m.Rs_AS > 0.0) && (xS_0 < 1.0);
  F_SA = loading(m, m.C_SA, theta, e_n, xS_0);
  kf_SA = (F_SA < F_SA_0) && (F_SA - m.Rs_SA < 0.0) && (xS_0 > 0.0);
  return kf_AS || kf_SA;
}

__host__ __device__ __forceinline__ ReturnMap return_map(const Mat& m, double theta, double e_n,
                                                double xS_0, double F_AS_0, double F_SA_0) {
#pragma clang fp contract(off)
  ReturnMap r;
  bool kf_AS, kf_SA;
  double F_AS, F_SA;
  flags_only(m, theta, e_n, xS_0, F_AS_0, F_SA_0, kf_AS, kf_SA, F_AS, F_SA);
  r.flags = (kf_AS ? FLAG_KF_AS : 0u) | (kf_SA ? FLAG_KF_SA : 0u);
  double xS = xS_0;
  if (kf_AS || kf_SA) {
    const double H_AS = kf_AS ? 1.0 : 0.0, H_SA = kf_SA ? 1.0 : 0.0;
    double F_AS_dot = F_AS - F_AS_0, F_SA_dot = F_SA - F_SA_0;
    double dh_AS = 0.0, dh_SA = 0.0;
    double RHS_AS = (F_AS - m.Rf_AS) * dh_AS + H_AS * F_AS_dot * (1.0 - xS);
    double RHS_SA = (F_SA - m.Rf_SA) * dh_SA - H_SA * F_SA_dot * xS;
    double resi = RHS_AS * RHS_AS + RHS_SA * RHS_SA;   // squared first (:403)
    int iter = 0;
    while (iter < 65 && resi > 1.0e-8) {
      ++iter;
      double m11 = -m.G_1 * dh_AS - (H_AS * F_AS_dot + H_AS * (1.0 - xS) * m.G_1) + (F_AS - m.Rf_AS);
      double m12 = -m.G_1 * dh_AS - (H_AS * F_AS_dot + H_AS * (1.0 - xS) * m.G_1);
      double m21 = -m.G_1 * dh_SA - (H_SA * F_SA_dot - H_SA * xS * m.G_1);
      double m22 = -m.G_1 * dh_SA - (H_SA * F_SA_dot - H_SA * xS * m.G_1) + (F_SA - m.Rf_SA);
      double det = m11 * m22 - m12 * m21;
      double r11 = m22 / det, r12 = -m12 / det, r21 = -m21 / det, r22 = m11 / det;
      double sol1 = r11 * RHS_AS + r12 * RHS_SA;
      double sol2 = r21 * RHS_AS + r22 * RHS_SA;
      dh_AS = dh_AS - sol1;
      dh_SA = dh_SA - sol2;
      xS = xS + dh_SA + dh_AS;                         // running totals (:429-432)
      F_AS = loading(m, m.C_AS, theta, e_n, xS);
      F_AS_dot = F_AS - F_AS_0;
      F_SA = loading(m, m.C_SA, theta, e_n, xS);
      F_SA_dot = F_SA - F_SA_0;
      RHS_AS = (F_AS - m.Rf_AS) * dh_AS + H_AS * F_AS_dot * (1.0 - xS);
      RHS_SA = (F_SA - m.Rf_SA) * dh_SA - H_SA * F_SA_dot * xS;
      resi = sqrt(RHS_AS * RHS_AS + RHS_SA * RHS_SA);
    }
    if (iter >= 65) {                                  // revert (:459-465)
      r.flags |= FLAG_NEWTON_FAIL;
      xS = xS_0;
      F_AS = F_AS_0;
      F_SA = F_SA_0;
    }
    r.flags |= (uint32_t)(iter & 0xff) << FLAG_ITER_SHIFT;
  }
  r.F_AS = F_AS;
  r.F_SA = F_SA;
  r.xS = fmax(fmin(1.0, xS), 0.0);                      // clamp after storing F's
  return r;
}

// Return map with the transformation update supplied by a surrogate: the
// branch flags are evaluated exactly (:342-374); on a flagged point xS takes
// the supplied value x_hat instead of the Newton iterate, and F_AS/F_SA are
// the loading functions at x_hat (stored before the clamp, as :484-498).
__host__ __device__ __forceinline__ ReturnMap return_map_given(const Mat& m, double theta, double e_n,
                                                               double xS_0, double F_AS_0, double F_SA_0,
                                                               double x_hat) {
#pragma clang fp contract(off)
  ReturnMap r;
  bool kf_AS, kf_SA;
  double F_AS, F_SA;
  flags_only(m, theta, e_n, xS_0, F_AS_0, F_SA_0, kf_AS, kf_SA, F_AS, F_SA);
  r.flags = (kf_AS ? FLAG_KF_AS : 0u) | (kf_SA ? FLAG_KF_SA : 0u);
  double xS = xS_0;
  if (kf_AS || kf_SA) {
    xS = x_hat;
    F_AS = loading(m, m.C_AS, theta, e_n, xS);
    F_SA = loading(m, m.C_SA, theta, e_n, xS);
  }
  r.F_AS = F_AS;
  r.F_SA = F_SA;
  r.xS = fmax(fmin(1.0, xS), 0.0);
  return r;
}

// Distance, in xS, from a candidate x to the return map's solution on a
// single-branch point.  With one branch active the reference's residual
// (:449-450) is linear in the increment dh = x - xS_0 (the dh^2 terms of
// (F(x) - R_f) dh and (F(x) - F_0)(1 - x) cancel because F is linear in x),
// so R(x) / R'(x) is exactly x minus the root, i.e. the size of the Newton
// correction the reference's first iteration would apply from x.  Returns a
// large value for two-branch points (always solved exactly) and when the
// slope vanishes.
__host__ __device__ __forceinline__ double consistency_residual(const Mat& m, double theta, double e_n,
                                                                double xS_0, double F_AS_0, double F_SA_0,
                                                                double x) {
  bool kf_AS, kf_SA;
  double F_AS, F_SA;
  flags_only(m, theta, e_n, xS_0, F_AS_0, F_SA_0, kf_AS, kf_SA, F_AS, F_SA);
  if (kf_AS == kf_SA) return 1e30;
  const double d = x - xS_0;
  // dF/dx of the loading function (eval_loading_function, beta = 1)
  const double k = -(2.0 * m.G * m.eL + 9.0 * m.alfa * m.alfa * m.K * m.eL);
  double R, dR;
  if (kf_AS) {
    const double Fx = loading(m, m.C_AS, theta, e_n, x);
    R = (Fx - m.Rf_AS) * d + (Fx - F_AS_0) * (1.0 - x);
    dR = k * d + (Fx - m.Rf_AS) + k * (1.0 - x) - (Fx - F_AS_0);
  } else {
    const double Fx = loading(m, m.C_SA, theta, e_n, x);
    R = (Fx - m.Rf_SA) * d - (Fx - F_SA_0) * x;
    dR = k * d + (Fx - m.Rf_SA) - k * x - (Fx - F_SA_0);
  }
  if (!(fabs(dR) > 0.0)) return 1e30;
  return fabs(R / dR);
}

// Kirchhoff stress (:487-491) as a symmetric tensor t00 t01 t02 t11 t12 t22.
template <typename R>
__host__ __device__ __forceinline__ void kirchhoff(const Mat& m, const Spectral<R>& s, double xS,
                                          R (&t)[6]) {
  R cvol = (R)(m.K * (s.theta - 3.0 * m.alfa * m.eL * xS));
  R exS = (R)(m.eL * xS);
  R twoG = (R)(2.0 * m.G);
  R c0 = twoG * (s.e[0] - exS * s.d[0]);
  R c1 = twoG * (s.e[1] - exS * s.d[1]);
  R c2 = twoG * (s.e[2] - exS * s.d[2]);
  const R* n = s.n;
  t[0] = cvol + c0 * n[0] * n[0] + c1 * n[3] * n[3] + c2 * n[6] * n[6];
  t[3] = cvol + c0 * n[1] * n[1] + c1 * n[4] * n[4] + c2 * n[7] * n[7];
  t[5] = cvol + c0 * n[2] * n[2] + c1 * n[5] * n[5] + c2 * n[8] * n[8];
  t[1] = c0 * n[0] * n[1] + c1 * n[3] * n[4] + c2 * n[6] * n[7];
  t[2] = c0 * n[0] * n[2] + c1 * n[3] * n[5] + c2 * n[6] * n[8];
  t[4] = c0 * n[1] * n[2] + c1 * n[4] * n[5] + c2 * n[7] * n[8];
}

// Column-major 3x3 inverse by cofactors (one reciprocal); returns det.  In
// FP32 the reciprocal is the hardware one (1 ulp).
template <typename R>
__host__ __device__ __forceinline__ R inv3(const R (&M)[9], R (&Mi)[9]) {
  R c00 = M[4] * M[8] - M[5] * M[7];
  R c01 = M[2] * M[7] - M[1] * M[8];
  R c02 = M[1] * M[5] - M[2] * M[4];
  R det = M[0] * c00 + M[3] * c01 + M[6] * c02;
  R r;
  if constexpr (sizeof(R) == 4) r = fast_rcp(det);
  else r = R(1) / det;
  Mi[0] = c00 * r;
  Mi[1] = c01 * r;
  Mi[2] = c02 * r;
  Mi[3] = (M[5] * M[6] - M[3] * M[8]) * r;
  Mi[4] = (M[0] * M[8] - M[2] * M[6]) * r;
  Mi[5] = (M[2] * M[3] - M[0] * M[5]) * r;
  Mi[6] = (M[3] * M[7] - M[4] * M[6]) * r;
  Mi[7] = (M[1] * M[6] - M[0] * M[7]) * r;
  Mi[8] = (M[0] * M[4] - M[1] * M[3]) * r;
  return det;
}

// S = F^-1 tau F^-T (col-major result) from Finv and symmetric tau.
template <typename R>
__host__ __device__ __forceinline__ void pk2(const R (&Fi)[9], const R (&t)[6], R (&S)[9]) {
  R T[9] = {t[0], t[1], t[2], t[1], t[3], t[4], t[2], t[4], t[5]};
  R A[9];  // A = Finv * tau
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      A[i + 3 * j] = Fi[i] * T[3 * j] + Fi[i + 3] * T[1 + 3 * j] + Fi[i + 6] * T[2 + 3 * j];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      S[i + 3 * j] = A[i] * Fi[j] + A[i + 3] * Fi[j + 3] + A[i + 6] * Fi[j + 6];
}

}  // namespace sma
