// HIP kernels of the explicit solid-mechanics step (gfx950, wave64).
//
// Per step, on one stream (captured into a hipGraph by the host):
//   sld_predict   nodal predictor, Dirichlet overwrite at t^{n+1}
//                 (mod_sld_solution_methods.f90:122-145 CD, :406-429 TW;
//                 sld_updbcs.f90:79-114 + sld_funbou.f90:95-136); clears
//                 f_int for the coloured scatter
//   sld_element   SoA gather into LDS element tiles (relative coordinates
//                 and displacements), reference Jacobian +
//                 deformation gradient at the 8 Gauss points, SMA return
//                 map, Kirchhoff stress, internal force
//                 (sld_element_operations.f90:119-256, elmca2.f90:105-111,
//                 sld_elmpre.f90:149-162, sld_builtin_materials.f90:316-371,
//                 sld_elmmat.f90:137-150, mod_matrix.f90:2327-2431).
//                 Two assemblies, both deterministic and bitwise equal:
//                 coloured scatter (one launch per colour, read-modify-write
//                 of f_int) or gather (one launch over all elements writing
//                 each corner force into its node-sorted slot; the
//                 corrector sums each node's contiguous slots in colour
//                 order, i.e. the same additions in the same order)
//   sld_face_load Neumann loads of boundary faces (sld_bouope.f90), when set
//   sld_correct   [gather f_int], f_ext (faces + nodal pattern + m g),
//                 a = (f_ext - f_int) / m, velocity corrector / TW Dirichlet
//                 zeroing (solric.f90:294-330; mod_sld_solution_methods.f90:
//                 181-210 CD, :494-527 TW); writes t^{n+1} = t^n + dt
//                 (setgts.f90:98); inside a captured run of steps it also
//                 runs the next step's predictor; optional energy sums
//                 (mod_sld_energy.f90) reduced by sld_energy_reduce
// The history (xS, F_AS, F_SA) and the clock live in two slots that swap
// roles every step (sld_updunk.f90:149-186, svegm TIME_N <- ITER_K): slot
// tn holds TIME_N / t^n, the step writes slot 1 - tn.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "sma_law.hpp"
#include "surrogate_features.hpp"

namespace sld {

// HEX08 natural derivatives at the 2x2x2 Gauss points: dN_a/dxi_k at GP g is
// 0.5 * s_ak * f(a,g,m1) * f(a,g,m2) where f = 0.5*(1 +- 0.577350269189626)
// depending on whether node a and GP g lie on the same side in direction m
// (mod_lagrange.f90:935-1000, mod_quadrature_gauss.f90:2320-2321,2372-2385).
__host__ __device__ __forceinline__ int node_sign(int a, int k) {
  // node order (-,-,-) (+,-,-) (+,+,-) (-,+,-) (-,-,+) (+,-,+) (+,+,+) (-,+,+)
  const int sx = 0x66, sy = 0xCC, sz = 0xF0;  // bit a set = +1
  int m = (k == 0) ? sx : (k == 1) ? sy : sz;
  return ((m >> a) & 1) ? 1 : -1;
}
__host__ __device__ __forceinline__ int gp_sign(int g, int k) {
  // GP order: xi outer, eta, zeta inner; bit = 1 means +0.5773...
  int bit = (k == 0) ? (g >> 2) & 1 : (k == 1) ? (g >> 1) & 1 : g & 1;
  return bit ? 1 : -1;
}
template <typename R>
__host__ __device__ __forceinline__ R hex_deriv(int a, int g, int k) {
  const double gpos = 0.577350269189626;
  const R cp = (R)(0.5 * (1.0 + gpos)), cm = (R)(0.5 * (1.0 - gpos));
  R f[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) f[m] = (node_sign(a, m) == gp_sign(g, m)) ? cp : cm;
  R s = (R)(0.5 * node_sign(a, k));
  if (k == 0) return s * f[1] * f[2];
  if (k == 1) return s * f[0] * f[2];
  return s * f[0] * f[1];
}
// Walsh-Hadamard form of the same derivatives.  With node signs s_am and GP
// signs s_gm (bit m of the index: xi = 4, eta = 2, zeta = 1), c the GP
// coordinate and (m1, m2) the two directions other than k,
//   dN_a/dxi_k(g) = 1/8 s_ak (1 + c s_a1 s_g1)(1 + c s_a2 s_g2),
// so sum_a x_a dN_a/dxi_k(g) = 1/8 [V(k) + c s_g1 V(k|m1) + c s_g2 V(k|m2)
// + c^2 s_g1 s_g2 V(7)] with V(h) = sum_a (prod_{m in h} s_am) x_a, and
// sum_g Q(g) dN_a/dxi_k(g) = 1/8 s_ak [U0 + c s_a1 U1 + c s_a2 U2
// + c^2 s_a1 s_a2 U12] with U the sums over g of Q(g), s_g1 Q(g), s_g2 Q(g),
// s_g1 s_g2 Q(g).  V and U are per element; only the signed weights differ
// between lanes.
constexpr double HEX_GPOS = 0.577350269189626;
// node with sign bits (bx, by, bz) in the reference node order
__host__ __device__ constexpr int hex_node(int bx, int by, int bz) {
  return 4 * bz + (by ? 3 - bx : bx);
}
// in-place Walsh-Hadamard butterfly over bit m of an 8-slot array: slot
// b | m becomes the signed difference (+ side minus - side)
template <typename R>
__device__ __forceinline__ void wht_stage(R (&v)[8], int m) {
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    if (b & m) continue;
    const R lo = v[b], hi = v[b | m];
    v[b] = lo + hi;
    v[b | m] = hi - lo;
  }
}

// U0, U1, U2, U12 of the 8 GP values q[b] of Q(3i+k): pair sums over the
// k-sign, then the two-direction Hadamard over (m1, m2).
template <typename R>
__device__ __forceinline__ void u4_of_q(const R* q, int k, R* u) {
  const int hk = 4 >> k, h1 = k == 0 ? 2 : 4, h2 = k == 2 ? 2 : 1;
  R p[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    // t = 2 b1 + b2: the GPs with m1-sign b1 and m2-sign b2
    const int b = ((t >> 1) ? h1 : 0) | ((t & 1) ? h2 : 0);
    p[t] = q[b] + q[b | hk];
  }
  const R s0 = p[0] + p[1], d0 = p[1] - p[0], s1 = p[2] + p[3], d1 = p[3] - p[2];
  u[0] = s0 + s1;   // sum
  u[1] = s1 - s0;   // m1-signed
  u[2] = d0 + d1;   // m2-signed
  u[3] = d1 - d0;   // m1 m2-signed
}

template <typename R>
__host__ __device__ __forceinline__ R hex_shape(int a, int g) {
  const double gpos = 0.577350269189626;
  const R cp = (R)(0.5 * (1.0 + gpos)), cm = (R)(0.5 * (1.0 - gpos));
  R f0 = (node_sign(a, 0) == gp_sign(g, 0)) ? cp : cm;
  R f1 = (node_sign(a, 1) == gp_sign(g, 1)) ? cp : cm;
  R f2 = (node_sign(a, 2) == gp_sign(g, 2)) ? cp : cm;
  return f0 * f1 * f2;
}

// Load tables: DISCRETE[,SMOOTH] (sld_funbou.f90 itask 2), rows (vx,vy,vz,t).
__device__ __forceinline__ double funbou(const double* __restrict__ tab, int npts,
                                         int smooth, int c, double t) {
#pragma clang fp contract(off)
  double x = 0.0;
  for (int i = 0; i < npts - 1; ++i) {
    double t1 = tab[4 * i + 3], t2 = tab[4 * (i + 1) + 3];
    if (t >= t1 && t <= t2) {
      double p1 = tab[4 * i + c], p2 = tab[4 * (i + 1) + c], m;
      if (smooth == 1) {
        m = (t - t1) / (t2 - t1);
        x = p1 + (p2 - p1) * (m * m * m) * (10.0 - 15.0 * m + 6.0 * (m * m));
      } else {
        m = (p2 - p1) / (t2 - t1);
        x = p1 + m * (t - t1);
      }
      break;
    }
  }
  double tlast = tab[4 * (npts - 1) + 3];
  if (t > tlast) x = tab[4 * (npts - 1) + c] * t / tlast;
  return x;
}

// Run health.  The reference stops the run on a non-positive Jacobian det F
// at a Gauss point (sld_elmpre.f90:229-240 keeps the smallest one and its
// element, sld_elmope.f90:106 -> sld_detjac.f90:93 calls runend) and on a
// NaN or infinite displacement or force residual (sld_cvgunk.f90:195-199).
// Here the kernel that meets such a state records it and sets `halted`;
// every step kernel returns at once while `halted` is set, so the device
// state stays as it was when the failure was found and the host reports it.
struct Health {
  // 0 = none; else (|det F| as FP32 bits) << 32 | (0xffffffff - element):
  // atomicMax keeps the most negative det F, ties to the lowest element
  unsigned long long detf;
  unsigned long long node;   // ~0 = none; else the lowest node with a non-finite value
  unsigned halted;           // bit 0: det F <= 0, bit 1: non-finite state
  unsigned pad;
};
constexpr unsigned HALT_DETF = 1u, HALT_NONFINITE = 2u;

__device__ __forceinline__ void report_detf(Health* h, int64_t e, double detF) {
  const float mag = fabsf((float)detF);
  const unsigned long long key = ((unsigned long long)__float_as_uint(mag) << 32) |
                                 (unsigned long long)(0xffffffffu - (uint32_t)e);
  atomicMax(&h->detf, key);
  atomicOr(&h->halted, HALT_DETF);
}
__device__ __forceinline__ void report_nonfinite(Health* h, int64_t node) {
  atomicMin(&h->node, (unsigned long long)node);
  atomicOr(&h->halted, HALT_NONFINITE);
}

// fixmask bits: 1, 2, 4 = prescribed x, y, z; FIX_FORCE = the node carries
// a nonzero entry of the nodal force pattern fext0 (set by the library).
constexpr int FIX_FORCE = 8;

struct NodalArgs {
  int64_t npoin;
  double* u[3];
  double* v[3];
  double* a[3];
  void* fint[3];               // double or float per precision mode
  int fint_fp32;
  const double* rmass;         // 1/m per node
  const uint8_t* fixmask;
  const int32_t* func_of_node; // may be null
  const double* func_table;    // [nfunc][maxpts][4]
  const int32_t* func_npts;
  const int32_t* func_smooth;
  int32_t func_maxpts;
  const double* fext0[3];      // may be null
  int32_t fext_func;           // -1 = none
  double* clock;               // [2] device time, slot tn = t^n
  int32_t tn;
  int32_t zero_fint;           // coloured scatter: the predictor clears f_int
  const int32_t* node_ptr;     // gather: [npoin+1] slot ranges per node, else null
  int32_t slot_width;          // gather with W slots per node (node i: W i ..), or 0
  const void* efor;            // gather: corner forces [slot][3] (double or float),
                               // slots of a node in ascending element order
  unsigned* sur_count;         // surrogate list segment counters, cleared per step
  // energy balance (mod_sld_energy.f90), when du is non-null: the predictor
  // stores du = u^{n+1} - u^n, the corrector writes per-block partial sums of
  // (f_int^{n+1} + f_int^n).du, (f_ext^{n+1} + f_ext^n).du and m v.v
  double* du[3];
  double* fint_prev[3];
  double* fext_prev[3];
  const double* mass;
  double* energy_part;         // [3][gridDim of the corrector]
  // boundary-face loads: [npoin+1] slot ranges per node (null = no faces)
  // and the face-corner forces of the step [slot][3]
  const int32_t* node_face_ptr;
  const double* fface;
  // gravity (sld_dev_set_gravity): f_ext += m g at every dof; mass is set
  int32_t gravity;
  double grav[3];
  // body force (sld_dev_set_body_force): f_ext += m b(t^{n+1}), b from load
  // function body_func by the rule of sld_funbou.f90 itask 3; mass is set
  int32_t body_func;           // -1 = none
  // Rayleigh damping (sld_dev_set_damping): the central-difference
  // predictor keeps v^n in vdamp (the element's veloc_sld(ITER_K)); the
  // corrector sums each node's corner damping forces (edamp, laid out as
  // efor) into the residual and stores them in fdamp at a run's last step
  double* vdamp[3];            // null: no damping or Tchamwa-Wielgosz
  const void* edamp;           // null: no damping
  double* fdamp[3];
  Health* health;              // run health (never null)
};

// Body-force value of component c at time t (sld_funbou.f90:142-167, itask
// 3): linear between the rows with t1 <= t < t2, value m t + b with
// m = (p2 - p1)/(t2 - t1), b = p1 - m t1; zero outside the table.
__device__ __forceinline__ double funbou_body(const double* __restrict__ tab, int npts, int c,
                                              double t) {
#pragma clang fp contract(off)
  for (int i = 0; i < npts - 1; ++i) {
    const double t1 = tab[4 * i + 3], t2 = tab[4 * (i + 1) + 3];
    if (t >= t1 && t < t2) {
      const double p1 = tab[4 * i + c], p2 = tab[4 * (i + 1) + c];
      const double m = (p2 - p1) / (t2 - t1);
      const double b = p1 - m * t1;
      return m * t + b;
    }
  }
  return 0.0;
}

// Boundary faces with Neumann loads (sld_bouope.f90), one QUA04 face per
// thread; corner forces go to node-sorted slots that the corrector sums.
struct FaceArgs {
  int64_t nface;
  const int32_t* lnodb;        // [nface][4]
  const int32_t* code;         // [nface] 3 pressure, 6 traction
  const double* value;         // [nface][3]
  const int32_t* func;         // [nface] or null
  int32_t transient, follower;
  const double* coord[3];
  const double* u[3];          // u^{n+1} (follower loads)
  const uint8_t* fixmask;
  const double* func_table;
  const int32_t* func_npts;
  const int32_t* func_smooth;
  int32_t func_maxpts;
  const int32_t* slot;         // [4 nface] slot of each face corner
  double* fface;               // [4 nface][3]
  const double* clock;
  int32_t tn;
};

// Predictor of node i from the step-n state (u, v, a): u^{n+1} with the
// Dirichlet values at t^{n+1} = cutim, and v^{n+1/2} (CD) or v^{n+1} (TW).
// scheme: 0 = central difference, 1 = Tchamwa-Wielgosz.  uvw = the state,
// already loaded.
__device__ __forceinline__ void predict_vals(const NodalArgs& A, int64_t i, int scheme, double dt,
                                             double phi, double cutim, int fix, const double u0[3],
                                             const double v0[3], const double a0[3], double un[3],
                                             double vn[3]) {
#pragma clang fp contract(off)
  const double t_end = cutim + dt, t_mid = 0.5 * (cutim + t_end);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double u = u0[d], v = v0[d], a = a0[d];
    if (scheme == 0) {
      vn[d] = v + (t_mid - cutim) * a;
      un[d] = u + dt * vn[d];
    } else {
      vn[d] = v + dt * a;
      un[d] = u + dt * v + phi * (dt * dt) * a;
    }
    if ((fix >> d) & 1) {
      un[d] = 0.0;
      const int fn = A.func_of_node ? A.func_of_node[i] : -1;
      if (fn >= 0)
        un[d] = funbou(A.func_table + (size_t)fn * A.func_maxpts * 4, A.func_npts[fn],
                       A.func_smooth[fn], d, cutim);
    }
  }
}

__device__ __forceinline__ void predict_node(const NodalArgs& A, int64_t i, int scheme, double dt,
                                             double phi, double cutim, double un[3], double vn[3]) {
  double u[3], v[3], a[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    u[d] = A.u[d][i];
    v[d] = A.v[d][i];
    a[d] = A.a[d][i];
  }
  predict_vals(A, i, scheme, dt, phi, cutim, A.fixmask[i], u, v, a, un, vn);
}

// Stores the predicted state (and du for the energy balance); the coloured
// scatter also clears f_int for the element launches that add into it.
__device__ __forceinline__ void store_prediction(const NodalArgs& A, int64_t i, const double u0[3],
                                                 const double un[3], const double vn[3],
                                                 const double vprev[3]) {
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (A.vdamp[0]) A.vdamp[d][i] = vprev[d];
    if (A.du[0]) A.du[d][i] = un[d] - u0[d];
    A.u[d][i] = un[d];
    A.v[d][i] = vn[d];
    if (A.zero_fint) {
      if (A.fint_fp32) ((float*)A.fint[d])[i] = 0.f;
      else ((double*)A.fint[d])[i] = 0.0;
    }
  }
}

template <int SCHEME>  // 0 = central difference, 1 = Tchamwa-Wielgosz
__global__ void __launch_bounds__(256) sld_predict(NodalArgs A, double dt, double phi) {
  if (A.health->halted) return;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.npoin) return;
  double u[3], v[3], a[3], un[3], vn[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    u[d] = A.u[d][i];
    v[d] = A.v[d][i];
    a[d] = A.a[d][i];
  }
  predict_vals(A, i, SCHEME, dt, phi, A.clock[A.tn] + dt, A.fixmask[i], u, v, a, un, vn);
  store_prediction(A, i, u, un, vn, v);
  if (!(isfinite(un[0]) && isfinite(un[1]) && isfinite(un[2]))) report_nonfinite(A.health, i);
}

// Gather assembly: f_int at node i = 0 + f_(e1) + f_(e2) + ... over the
// element corners at the node in ascending (colour-sorted) element order,
// which is the order in which the coloured scatter adds them.
// Slots padded to W per node (W = the mesh's largest corner count, unused
// slots zero): node i's values sit at a fixed place, so the loads need no
// slot range and the wave's loads are contiguous; W = 8 (hex meshes) loads
// all 24 values at once.  Trailing zeros leave the ordered sum unchanged.
// Without padding, a node with 8 corners still loads them in one round trip.
template <typename R>
__device__ __forceinline__ void gather_fint(const NodalArgs& A, int64_t i, double fi[3],
                                            bool store, const void* slots = nullptr) {
  const int W = A.slot_width;
  const int32_t k0 = W ? W * (int32_t)i : A.node_ptr[i];
  const int32_t k1 = W ? k0 + W : A.node_ptr[i + 1];
  const R* ef = (const R*)(slots ? slots : A.efor) + 3 * (int64_t)k0;
  R s0 = R(0), s1 = R(0), s2 = R(0);
  if (k1 - k0 == 8) {
    R c[24];
#pragma unroll
    for (int j = 0; j < 24; ++j) c[j] = ef[j];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      s0 += c[3 * j];
      s1 += c[3 * j + 1];
      s2 += c[3 * j + 2];
    }
  } else {
    for (int32_t k = 0; k < k1 - k0; ++k) {
      s0 += ef[3 * k];
      s1 += ef[3 * k + 1];
      s2 += ef[3 * k + 2];
    }
  }
  if (store && !slots) {
    ((R*)A.fint[0])[i] = s0;
    ((R*)A.fint[1])[i] = s1;
    ((R*)A.fint[2])[i] = s2;
  }
  fi[0] = (double)s0;
  fi[1] = (double)s1;
  fi[2] = (double)s2;
}

// Per-node energy terms of one step: internal-force and external-force work
// over the displacement increment (trapezoidal, mod_sld_energy.f90:113-131)
// and m v.v (mod_sld_energy.f90:134-140); f_int^{n+1} and f_ext^{n+1} replace
// the previous step's.
__device__ __forceinline__ void node_energy(const NodalArgs& A, int64_t i, const double fg[3],
                                            const double fe[3], const double vel[3],
                                            double e[3]) {
#pragma clang fp contract(off)
  double ie = 0.0, wk = 0.0, ke = 0.0;
  const double m = A.mass[i];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double du = A.du[d][i];
    ie = ie + (fg[d] + A.fint_prev[d][i]) * du;
    wk = wk + (fe[d] + A.fext_prev[d][i]) * du;
    A.fext_prev[d][i] = fe[d];
    const double v = vel[d];
    ke = ke + v * v * m;
    A.fint_prev[d][i] = fg[d];
  }
  e[0] = ie;
  e[1] = wk;
  e[2] = ke;
}

// Corrector of node i.  FUSE (inside a captured run of steps): the
// corrected a and v stay in registers and the next step's predictor runs
// on them at t^{n+2} = t^{n+1} + dt, with the same arithmetic as its own
// kernel; a, v^{n+1} and f_int reach memory only at the last step of the run.
// The node's loads are all issued first; the load function of the nodal
// force pattern is evaluated only at nodes that carry one (FIX_FORCE).
template <int SCHEME, bool FUSE>
__device__ __forceinline__ void correct_node(const NodalArgs& A, int64_t i, double dt, double phi,
                                             double e[3]) {
#pragma clang fp contract(off)
  const double rm = A.rmass[i];
  const int fix = A.fixmask[i];
  double fg[3];
  if (A.node_ptr) {
    if (A.fint_fp32) gather_fint<float>(A, i, fg, !FUSE);
    else gather_fint<double>(A, i, fg, !FUSE);
  } else {
#pragma unroll
    for (int d = 0; d < 3; ++d)
      fg[d] = A.fint_fp32 ? (double)((const float*)A.fint[d])[i] : ((const double*)A.fint[d])[i];
  }
  double v0[3], u0[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    v0[d] = A.v[d][i];
    if (FUSE) u0[d] = A.u[d][i];
  }
  const double cutim = A.clock[A.tn] + dt;  // t^{n+1}
  const double t_end = cutim + dt, t_mid = 0.5 * (cutim + t_end);
  // f_ext: face loads (summed per node in face order) + fx fext0
  double fe[3] = {0.0, 0.0, 0.0};
  if (A.node_face_ptr) {
    const int32_t k0 = A.node_face_ptr[i], k1 = A.node_face_ptr[i + 1];
    for (int32_t k = k0; k < k1; ++k) {
#pragma unroll
      for (int d = 0; d < 3; ++d) fe[d] = fe[d] + A.fface[3 * (int64_t)k + d];
    }
  }
  if (A.fext_func >= 0 && (fix & FIX_FORCE)) {
    const double fx = funbou(A.func_table + (size_t)A.fext_func * A.func_maxpts * 4,
                             A.func_npts[A.fext_func], A.func_smooth[A.fext_func], 0, cutim);
#pragma unroll
    for (int d = 0; d < 3; ++d) fe[d] = fe[d] + fx * A.fext0[d][i];
  }
  if (A.gravity || A.body_func >= 0) {
    // the element load sum_g N_a rho (g + b) w detJ (sld_elmmat.f90:152-157)
    // of every element at the node: the lumped mass times g + b(t^{n+1})
    const double m = A.mass[i];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      double gb = A.gravity ? A.grav[d] : 0.0;
      if (A.body_func >= 0)
        gb = gb + funbou_body(A.func_table + (size_t)A.body_func * A.func_maxpts * 4,
                              A.func_npts[A.body_func], d, cutim);
      fe[d] = fe[d] + m * gb;
    }
  }
  // Rayleigh damping forces of the node's element corners
  double fd[3] = {0.0, 0.0, 0.0};
  if (A.edamp) {
    if (A.fint_fp32) gather_fint<float>(A, i, fd, false, A.edamp);
    else gather_fint<double>(A, i, fd, false, A.edamp);
    if (!FUSE) {
#pragma unroll
      for (int d = 0; d < 3; ++d) A.fdamp[d][i] = fd[d];
    }
  }
  double an[3], vn[3];
  bool finite = true;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double fi = fg[d];
    double r = 0.0 - fi;
    r = r - fd[d];
    r = r + fe[d];
    finite = finite && isfinite(r);
    double acc = 0.0 + r * rm;
    if (SCHEME == 0) {
      an[d] = acc;
      vn[d] = v0[d] + (t_end - t_mid) * acc;
    } else {
      vn[d] = v0[d];
      if ((fix >> d) & 1) { vn[d] = 0.0; acc = 0.0; }
      an[d] = acc;
    }
    if (!FUSE) {
      A.a[d][i] = an[d];
      if (SCHEME == 0 || ((fix >> d) & 1)) A.v[d][i] = vn[d];
    }
  }
  if (A.energy_part) node_energy(A, i, fg, fe, vn, e);
  if (!finite) report_nonfinite(A.health, i);
  if (FUSE) {
    double un[3], vp[3];
    predict_vals(A, i, SCHEME, dt, phi, cutim + dt, fix, u0, vn, an, un, vp);
    store_prediction(A, i, u0, un, vp, vn);
    if (!(isfinite(un[0]) && isfinite(un[1]) && isfinite(un[2]))) report_nonfinite(A.health, i);
  }
  if (i == 0) {
    A.clock[1 - A.tn] = cutim;
  }
}

// Fixed-order block sum of three values; thread 0 writes part[k*gridDim.x + b].
__device__ __forceinline__ void block_sum3(double e[3], double* part) {
  __shared__ double red[3][256];
  const int t = threadIdx.x;
#pragma unroll
  for (int k = 0; k < 3; ++k) red[k][t] = e[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
#pragma unroll
      for (int k = 0; k < 3; ++k) red[k][t] = red[k][t] + red[k][t + w];
    }
    __syncthreads();
  }
  if (t == 0) {
#pragma unroll
    for (int k = 0; k < 3; ++k) part[(size_t)k * gridDim.x + blockIdx.x] = red[k][0];
  }
}

template <int SCHEME, bool FUSE>
__global__ void __launch_bounds__(256) sld_correct(NodalArgs A, double dt, double phi) {
#pragma clang fp contract(off)
  if (A.health->halted) return;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double e[3] = {0.0, 0.0, 0.0};
  if (i < A.npoin) correct_node<SCHEME, FUSE>(A, i, dt, phi, e);
  // the next step's surrogate list starts empty
  if (A.sur_count)
    for (int64_t k = i; k < mlp::LIST_SEGS; k += (int64_t)gridDim.x * blockDim.x)
      A.sur_count[k * mlp::LIST_STRIDE] = 0u;
  if (A.energy_part) block_sum3(e, A.energy_part);
}

// Sums the corrector's partials in a fixed order and advances the totals
// acc = {allie, allwk, allke, etota, steps} (mod_sld_energy.f90:141-162): the
// first step after a reset takes the full increment, later ones half the
// trapezoid.
__global__ void __launch_bounds__(256) sld_energy_reduce(const double* part, int nparts,
                                                         double* acc, const Health* health) {
#pragma clang fp contract(off)
  if (health->halted) return;
  double e[3] = {0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nparts; b += 256) {
#pragma unroll
    for (int k = 0; k < 3; ++k) e[k] = e[k] + part[(size_t)k * nparts + b];
  }
  __shared__ double red[3][256];
  const int t = threadIdx.x;
#pragma unroll
  for (int k = 0; k < 3; ++k) red[k][t] = e[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
#pragma unroll
      for (int k = 0; k < 3; ++k) red[k][t] = red[k][t] + red[k][t + w];
    }
    __syncthreads();
  }
  if (t == 0) {
    const double w = acc[4] == 0.0 ? 1.0 : 0.5;
    acc[0] = acc[0] + w * fabs(red[0][0]);
    acc[1] = acc[1] + w * fabs(red[1][0]);
    acc[2] = 0.5 * red[2][0];
    acc[3] = acc[0] + acc[2] - acc[1];
    acc[4] = acc[4] + 1.0;
  }
}

// Face loads of one step at t^{n+1} (sld_bouope.f90:88-277 with the values
// of sld_updbcs.f90:153-181): geometry of the face's 4 nodes (reference, or
// current with follower loads); value = the constant one, or under transient
// conditions (column x of the face's function, 0, 0) and 0 without one
// (sld_funbou.f90 itask 1); at the 2x2 Gauss points (s outer, t inner,
// mod_quadrature_gauss.f90:2361-2370) the tangents of bouder
// (mod_bouder.f90:74-89), n = t_s x t_t, |n|, the unit normal of chenor, the
// traction (code 3: value_1 n/|n|, code 6: value) and the corner forces
// |n| tract N_a on dofs that are not prescribed (QUA04 of mod_lagrange.f90).
__global__ void __launch_bounds__(256) sld_face_load(FaceArgs F, double dt) {
#pragma clang fp contract(off)
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= F.nface) return;
  const double cutim = F.clock[F.tn] + dt;
  int32_t nb[4];
  double x[4][3];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    nb[a] = F.lnodb[4 * f + a];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      x[a][d] = F.coord[d][nb[a]];
      if (F.follower) x[a][d] = x[a][d] + F.u[d][nb[a]];
    }
  }
  double val[3];
  if (F.transient) {
    const int fn = F.func ? F.func[f] : -1;
    val[0] = fn >= 0 ? funbou(F.func_table + (size_t)fn * F.func_maxpts * 4, F.func_npts[fn],
                              F.func_smooth[fn], 0, cutim)
                     : 0.0;
    val[1] = 0.0;
    val[2] = 0.0;
  } else {
#pragma unroll
    for (int d = 0; d < 3; ++d) val[d] = F.value[3 * f + d];
  }
  const int code = F.code[f];
  const double g = 0.577350269189626;
  double elfex[4][3];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int d = 0; d < 3; ++d) elfex[a][d] = 0.0;
  for (int ig = 0; ig < 4; ++ig) {
    const double s = ig < 2 ? -g : g, t = (ig & 1) ? g : -g, st = s * t;
    const double N[4] = {(1.0 - t - s + st) * 0.25, (1.0 - t + s - st) * 0.25,
                         (1.0 + t + s + st) * 0.25, (1.0 + t - s - st) * 0.25};
    const double ds[4] = {(-1.0 + t) * 0.25, (1.0 - t) * 0.25, (1.0 + t) * 0.25,
                          (-1.0 - t) * 0.25};
    const double dtt[4] = {(-1.0 + s) * 0.25, (-1.0 - s) * 0.25, (1.0 + s) * 0.25,
                           (1.0 - s) * 0.25};
    double t1[3], t2[3], n[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      t1[d] = 0.0;
      t2[d] = 0.0;
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        t1[d] = t1[d] + x[a][d] * ds[a];
        t2[d] = t2[d] + x[a][d] * dtt[a];
      }
    }
    n[0] = t1[1] * t2[2] - t1[2] * t2[1];
    n[1] = t1[2] * t2[0] - t1[0] * t2[2];
    n[2] = t1[0] * t2[1] - t1[1] * t2[0];
    const double eucta = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    const double gbsur = 1.0 * eucta;
    double tract[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) tract[d] = code == 3 ? (n[d] / eucta) * val[0] : val[d];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const int fix = F.fixmask[nb[a]];
#pragma unroll
      for (int d = 0; d < 3; ++d)
        if (!((fix >> d) & 1)) elfex[a][d] = elfex[a][d] + gbsur * tract[d] * N[a];
    }
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int64_t k = F.slot[4 * f + a];
#pragma unroll
    for (int d = 0; d < 3; ++d) F.fface[3 * k + d] = elfex[a][d];
  }
}

// ----------------------------------------------------------------------------
// Element kernel
// ----------------------------------------------------------------------------
struct ElemArgs {
  int32_t e_begin, e_end;       // range of the launch in elem_list (or of elements)
  const int32_t* elem_list;     // scatter: elements sorted by colour; gather: null (identity)
  int64_t nelem;
  const int32_t* lnods;         // [nelem][8]
  const double* coord[3];       // SoA reference coordinates
  const double* u[3];           // SoA displacement (ITER_K)
  const double* props;          // [nmat][18]
  const double* thresholds;     // [nmat][4] mat_thresholds of each material
  const int32_t* mat_of_elem;   // [nelem] or null
  const double* E_of_elem;      // [nelem] or null
  const double* kg;             // [nelem][2] bulk and shear modulus per element
  const uint8_t* law_of_elem;   // [nelem] 1 = neo-Hookean (model 102), 0 = SMA; null = all SMA
  const double* sdv_n[3];       // TIME_N history, SoA over GP = 8*e + g
  double* sdv_k[3];             // ITER_K history
  void* fint[3];                // scatter: nodal internal force (double or float)
  void* efor;                   // gather: corner forces [slot][3] (double or float)
  const int32_t* corner_slot;   // gather: [8 * e + a] -> slot of the corner at its node
  double* S_out;                // optional PK2 [9][ngp]
  double* P_out;                // optional first Piola F S [9][ngp] (with S_out)
  uint32_t* flags_out;          // optional [ngp]
  unsigned long long* counters; // [COUNTER_SLOTS][8] per-flag Gauss-point counts
  Health* health;               // run health (never null)
  // Rayleigh damping (DAMP kernels): veloc_sld(ITER_K) of the step (v^n in
  // the central difference, the predicted v^{n+1} in Tchamwa-Wielgosz),
  // (alpha, beta) and density per material, corner damping forces
  const double* vel[3];
  const double* damping;        // [nmat][2]
  const double* rho;            // [nmat]
  void* edamp;                  // [slot][3] (double or float)
  int32_t xcd_remap;            // 1 = spread consecutive tiles per XCD
  int32_t tiles_per_block;      // > 0: persistent launch, tiles per workgroup
  // surrogate (mode 2): per-GP status from the prepass (0 not flagged,
  // 1 listed, 2 flagged and not listed), the listed points' slots, their
  // features and the network's outputs by slot
  uint8_t* sur_status;
  int32_t* sur_slot;            // [gp], valid where status is 1
  float* sur_feat;              // [cap][16]
  const float* sur_Y;           // [cap][12]
  double sur_tol;               // acceptance tolerance on xS
  unsigned* sur_count;          // the list's segment counters (mlp::LIST_SEGS)
  int64_t sur_seg_cap;          // slots per segment
  double T_feat;                // temperature fed to the surrogate (inert in the law)
  // surrogate candidates: per element, 1 when one of its points transformed
  // in the previous step (the prepass lists points of those elements only;
  // every other point takes the exact map), and the flags the step writes
  // for the next one; null: every element is a candidate
  const uint8_t* sur_cand;
  uint8_t* sur_cand_next;
};

// Flag counters live in COUNTER_SLOTS slots of 8 (one 64-byte line each);
// workgroup b adds into slot b % COUNTER_SLOTS, so the per-wave atomics of a
// launch spread over many addresses instead of serialising on one.  The host
// sums the slots.
constexpr int COUNTER_SLOTS = 256;

// Surrogate mode, one Gauss point: a listed point takes the network's xS
// when it lies within tol of the return map's solution (|R(x)/R'(x)| of the
// single-branch consistency residual, exact because that residual is linear
// in the increment on one branch) and in [-1e-3, 1 + 1e-3]; every other
// point takes the exact map.  Flags: FLAG_SURR_USED on accepted transforming
// points, FLAG_SURR_FALLBACK on flagged points that took the exact map.
__device__ inline sma::ReturnMap surrogate_or_exact(const ElemArgs& A, const sma::Mat& m, int64_t gp,
                                                   double theta, double e_n, double xS0,
                                                   double FAS0, double FSA0) {
  const uint8_t st = A.sur_status[gp];
  bool ok = false;
  double x = 0.0;
  if (st == 1) {
    const int64_t slot = A.sur_slot[gp];
    x = xS0 + (double)A.sur_Y[slot * mlp::NOUT + 11] / mlp::Scale::dx;
    ok = (x >= -1e-3) && (x <= 1.0 + 1e-3) &&
         sma::consistency_residual(m, theta, e_n, xS0, FAS0, FSA0, x) <= A.sur_tol;
  }
  sma::ReturnMap rm;
  if (ok) {
    rm = sma::return_map_given(m, theta, e_n, xS0, FAS0, FSA0, x);
    if (rm.flags & 3u) rm.flags |= sma::FLAG_SURR_USED;
  } else {
    rm = sma::return_map(m, theta, e_n, xS0, FAS0, FSA0);
    if (st != 0 && (rm.flags & 3u)) rm.flags |= sma::FLAG_SURR_FALLBACK;
  }
  return rm;
}

// Wave-level flag counters: one 64-bit atomic per wave and per set bit type.
__device__ __forceinline__ void count_flags(unsigned long long* ctr, uint32_t fl, bool valid) {
  if (__ballot(valid && fl != 0u) == 0ull) return;   // wave-uniform: nothing to count
  ctr += 8 * (blockIdx.x % COUNTER_SLOTS);
  const uint32_t bits[7] = {sma::FLAG_KF_AS, sma::FLAG_KF_SA, sma::FLAG_NEWTON_FAIL,
                            sma::FLAG_EIGEN_FAIL, sma::FLAG_DETF_NONPOS,
                            sma::FLAG_SURR_USED, sma::FLAG_SURR_FALLBACK};
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int b = 0; b < 7; ++b) {
    unsigned long long m = __ballot(valid && (fl & bits[b]));
    if (m && lane == 0) atomicAdd(ctr + b, (unsigned long long)__popcll(m));
  }
}

// EPB elements per block, 8 threads (one per Gauss point) per element.
// R = kinematics/stress precision; the return map is always FP64.
// SUR = 0 plain; 1 surrogate prepass (flags + features of transforming
// points, no stress, no scatter); 2 element step using accepted surrogate
// answers (exact return map otherwise).
// GATHER = false: read-modify-write of f_int (the launch covers one colour);
// true: per-corner forces to efor (the launch may cover every element).
#ifndef SLD_WAVES_F64
#define SLD_WAVES_F64 3
#endif
#ifndef SLD_WAVES_F32
#define SLD_WAVES_F32 4
#endif
#ifndef SLD_WAVES_F64_SMA
#define SLD_WAVES_F64_SMA 4
#endif
// LAWS = 0: each element takes its own law (law_of_elem); LAWS = 1: every
// element of the launch is SMA (a mesh without neo-Hookean elements, or the
// SMA range of a law-grouped mesh), without the neo-Hookean path, whose
// registers would otherwise add to the SMA path's (FP64: 124 VGPRs and 4
// waves per SIMD against 163 and 3); LAWS = 2: every element of the launch
// is neo-Hookean (the tissue range of a law-grouped mesh), without the SMA
// path (FP64: 128 VGPRs, 4 waves per SIMD, a few values spilled; it runs 5%
// faster than at 3 waves without spills; FP32: 64 VGPRs, 8 waves, 4.3 KB of
// LDS per workgroup).  Kernels that also store the stresses (STORE) keep 3
// waves in FP64 and 6 in the FP32 neo-Hookean kernel, where more would spill.
template <typename R, int LAWS = 0, bool DAMP = false, bool STORE = false>
struct ElemOccupancy {
  static constexpr int waves = DAMP ? (sizeof(R) == 8 ? 1 : 2)
                               : LAWS == 2 ? (sizeof(R) == 8 ? 4 : (STORE ? 6 : 8))
                               : LAWS == 1 ? (sizeof(R) == 8 ? (STORE ? 3 : SLD_WAVES_F64_SMA)
                                                             : SLD_WAVES_F32)
                                           : (sizeof(R) == 8 ? SLD_WAVES_F64 : SLD_WAVES_F32);
};

// Elements of R in one element's LDS block, padded from n: a multiple of 16
// bytes whose size in 4-byte words is an odd multiple of 4, so that the
// blocks of 8 consecutive elements start 4 banks apart (mod 32) and values
// read at the same offset by the 8 elements of a wave fall in 8 different
// 4-bank groups.
constexpr int lds_block(int n, int rsize) {
  int m = n;
  while ((m * rsize) % 16 != 0 || ((m * rsize) / 4) % 8 != 4) ++m;
  return m;
}

// c with its sign flipped when neg is 1 (neg is 0 or 1): an integer xor on
// the sign bit, so that a lane-dependent signed weight costs no pair of
// constant registers held through the kernel
template <typename R>
__device__ __forceinline__ R sign_if(R c, int neg) {
  if constexpr (sizeof(R) == 8)
    return __longlong_as_double(__double_as_longlong(c) ^ ((long long)(neg & 1) << 63));
  else
    return __int_as_float(__float_as_int(c) ^ (int)((unsigned)(neg & 1) << 31));
}

// The value of x, opaque to the optimiser: quantities derived from the lane
// index (shape-function derivatives at the lane's Gauss point or node) are
// then recomputed where they are used instead of being hoisted out of the
// tile loop and held in registers through the constitutive update.
__device__ __forceinline__ int opaque(int x) {
  asm volatile("" : "+v"(x));
  return x;
}

// DAMP: Rayleigh damping (sld_elmmat.f90:160-163, :174-194) -- the
// element also gathers the velocity, forms at each Gauss point the
// consistent-mass term q = alpha rho detJ v(g) and the beta term's
// z_k = detJ sum_j J^-1_kj w_j, w_j = sum_k L_kj S_jk (L the velocity
// gradient, S the PK2 stress), and writes per corner the damping force
// alpha sum_g N_a q(g) - sum_g sum_k dN_a/dxi_k z_k to edamp.
template <typename R, int EPB, bool STORE, int SUR = 0, bool GATHER = false, int LAWS = 0,
          bool DAMP = false>
__global__ void __launch_bounds__(EPB * 8, (ElemOccupancy<R, LAWS, DAMP, STORE>::waves))
    sld_element(ElemArgs A) {
  constexpr int NC = DAMP ? 9 : 6;        // gathered components: x, u (, v)
  constexpr int NQ = DAMP ? 15 : 9;
  // The relative nodal values are dead once V(h) is formed (a barrier
  // follows), before the Gauss-point rows are first written: one LDS block
  // serves both (the barrier at the end of a tile keeps the next tile's
  // gather off the rows phase 3 still reads).  Each element's block is
  // padded (lds_block) so that the 8 elements of a wave start in different
  // LDS banks: the eight lanes of an element read the same value at once,
  // and unpadded power-of-two strides put all eight elements' values in one
  // bank.
  // The rows of the relative values and of the Gauss-point values have 9
  // entries (18 banks apart in FP64), so that the 8 lanes of an element
  // reading 8 different rows hit different banks, and FP64 U rows 5.
  // Without damping, V(h) is dead once the kinematics are formed, and the U
  // sums (and the FP64 det J kept through the law) take its block.
  constexpr bool U4_IN_V = !DAMP;
  constexpr int ROW = 9;
  constexpr int U4S = (!DAMP && sizeof(R) == 8) ? 5 : 4;
  constexpr int VC = DAMP ? 10 : ((LAWS == 2 || sizeof(R) == 8) ? 6 : 8);  // V(h) row
  constexpr int U4R = DAMP ? 11 : 8;
  constexpr int RELN = lds_block(NC * ROW, sizeof(R));
  constexpr int QN = lds_block(NQ * ROW, sizeof(R));
  constexpr int VN = lds_block(U4_IN_V ? (8 * VC > U4R * U4S ? 8 * VC : U4R * U4S) : 8 * VC,
                               sizeof(R));
  constexpr int UN = U4_IN_V ? 1 : lds_block(U4R * U4S, sizeof(R));
  __shared__ union {
    R rel[EPB][RELN];                     // [c][a]: x_a - x_0 (0..2), u_a - u_0 (3..5), v_a - v_0 (6..8)
    R q[EPB][QN];                         // [j][g]: Q(3i+k) (, q_i, z_k) at each Gauss point
  } s_relq;
  // V(h) of the components ([h][c]; FP32 rows of 8 for aligned reads, FP64
  // rows of 6 are 16-byte aligned already)
  __shared__ R s_vm[EPB][VN];
  __shared__ R s_u4m[U4_IN_V ? 1 : EPB][UN];  // [j][0..3]: U0, U1, U2, U12 of Q(0..7) (, z_0..2)
  auto s_rel = [&](int l) { return reinterpret_cast<R(*)[ROW]>(&s_relq.rel[l][0]); };
  auto s_q = [&](int l) { return reinterpret_cast<R(*)[ROW]>(&s_relq.q[l][0]); };
  auto s_v = [&](int l) { return reinterpret_cast<R(*)[VC]>(&s_vm[l][0]); };
  auto s_u4 = [&](int l) {
    if constexpr (U4_IN_V) return reinterpret_cast<R(*)[U4S]>(&s_vm[l][0]);
    else return reinterpret_cast<R(*)[U4S]>(&s_u4m[l][0]);
  };
  __shared__ R s_wq[DAMP ? EPB : 1][3][8];   // Walsh-Hadamard coefficients of q_i
  __shared__ double s_v0[DAMP ? EPB : 1][3]; // node 0's velocity

  if (A.health->halted) return;
  const int tid = threadIdx.x;
  const int le = tid >> 3, g = tid & 7;
  int bid = blockIdx.x;
  if (A.xcd_remap) {
    // workgroups are dealt round-robin to the 8 XCDs: workgroup b runs on
    // XCD b % 8 as that XCD's (b / 8)-th.  Give each XCD a contiguous run of
    // tiles so neighbouring elements share its L2: XCD x owns q + (x < r)
    // workgroups (nb = 8 q + r), its run starts at x q + min(x, r).
    const int nb = gridDim.x, q = nb >> 3, r = nb & 7;
    const int xcd = bid & 7, k = bid >> 3;
    bid = xcd * q + min(xcd, r) + k;
  }
  // persistent launches give each workgroup a contiguous run of tiles
  const int ntiles = (A.e_end - A.e_begin + EPB - 1) / EPB;
  const int per = A.tiles_per_block > 0 ? A.tiles_per_block : 1;
  const int t_begin = bid * per, t_end = min(ntiles, t_begin + per);
  for (int blk = t_begin; blk < t_end; ++blk) {
  const int i_el = A.e_begin + blk * EPB + le;
  const bool valid = i_el < A.e_end;
  const int e = !valid ? 0 : A.elem_list ? A.elem_list[i_el] : i_el;
  if constexpr (SUR == 1) {
    // prepass: a workgroup without a candidate element only marks its
    // points as not listed (workgroup-uniform, before any barrier)
    if (A.sur_cand && !__syncthreads_or(valid && A.sur_cand[e])) {
      if (valid) A.sur_status[(int64_t)e * 8 + g] = 0;
      continue;
    }
  }

  // -- phase 1: gather the element's 8 nodes (thread g loads node g); node
  // 0's values come from the element's first lane by a wave shuffle
  int node = 0;
  double xa[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) xa[c] = 0.0;
  if (valid) {
    node = A.lnods[(int64_t)e * 8 + g];
    xa[0] = A.coord[0][node];
    xa[1] = A.coord[1][node];
    xa[2] = A.coord[2][node];
    xa[3] = A.u[0][node];
    xa[4] = A.u[1][node];
    xa[5] = A.u[2][node];
    if constexpr (DAMP) {
      xa[6] = A.vel[0][node];
      xa[7] = A.vel[1][node];
      xa[8] = A.vel[2][node];
    }
  }
  {
    const int lane0 = (tid & 63) & ~7;   // lane of the element's node 0
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const double x0 = __shfl(xa[c], lane0, 64);
      if (valid) s_rel(le)[c][g] = (R)(xa[c] - x0);
      if constexpr (DAMP)
        if (c >= 6 && valid && g == 0) s_v0[le][c - 6] = x0;
    }
  }
  __syncthreads();
  // thread c: V(h) of component c, nodes taken in sign-bit order
  if (valid && g < NC) {
    for (int c = opaque(g); c < NC; c += 8) {
      R v[8];
#pragma unroll
      for (int b = 0; b < 8; ++b) v[b] = s_rel(le)[c][hex_node((b >> 2) & 1, (b >> 1) & 1, b & 1)];
      wht_stage(v, 4);
      wht_stage(v, 2);
      wht_stage(v, 1);
#pragma unroll
      for (int h = 0; h < 8; ++h) s_v(le)[h][c] = v[h];
    }
  }
  __syncthreads();

  // -- phase 2: Gauss point g of element e
  uint32_t fl = 0;
  if (valid) {
    const int gk = opaque(g);
    // signed weights of the lane's GP: c s_gm / 8 and c^2 s_gm s_gn / 8;
    // row i of J (i < 3) or of G_u (i = 3..5) at the GP, column k
    auto grad_row = [&](int i, R (&out)[3]) {
      const R c1 = (R)(0.125 * HEX_GPOS), c2 = (R)(0.125 * HEX_GPOS * HEX_GPOS);
      const R k0 = (R)0.125;
      const R ax = sign_if(c1, ((gk >> 2) & 1) ^ 1), ay = sign_if(c1, ((gk >> 1) & 1) ^ 1),
              az = sign_if(c1, (gk & 1) ^ 1);
      const R byz = sign_if(c2, ((gk >> 1) ^ gk) & 1);
      const R bxz = sign_if(c2, ((gk >> 2) ^ gk) & 1);
      const R bxy = sign_if(c2, ((gk >> 2) ^ (gk >> 1)) & 1);
      const R* V7 = s_v(le)[7];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int hk = 4 >> k, h1 = k == 0 ? 2 : 4, h2 = k == 2 ? 2 : 1;
        const R w1 = k == 0 ? ay : ax, w2 = k == 2 ? ay : az, w3 = k == 0 ? byz : k == 1 ? bxz : bxy;
        out[k] = k0 * s_v(le)[hk][i] + w1 * s_v(le)[hk | h1][i] + w2 * s_v(le)[hk | h2][i] +
                 w3 * V7[i];
      }
    };
    // J^-1 first (J is then dead), H = F - I = G_u J^-1 (G_u row by row)
    R Ji[9];
    R detJ;
    {
      R J[9];  // col-major: J(i,k) = J[i + 3k]
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        R r3[3];
        grad_row(i, r3);
#pragma unroll
        for (int k = 0; k < 3; ++k) J[i + 3 * k] = r3[k];
      }
      detJ = sma::inv3<R>(J, Ji);
    }
    R H[9];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      R gu[3];
      grad_row(3 + i, gu);
#pragma unroll
      for (int j = 0; j < 3; ++j)
        H[i + 3 * j] = gu[0] * Ji[3 * j] + gu[1] * Ji[1 + 3 * j] + gu[2] * Ji[2 + 3 * j];
    }
    // current Jacobian j = J + G_u = F J: j^-1 = J^-1 F^-1 and det j =
    // det J det F, formed before the constitutive update
    R jci[9];
    R detF;
    {
      R Fm1[9], Fi[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) Fm1[q] = H[q] + ((q % 4 == 0) ? R(1) : R(0));
      detF = sma::inv3<R>(Fm1, Fi);
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
          jci[r + 3 * c] = Ji[r] * Fi[3 * c] + Ji[r + 3] * Fi[1 + 3 * c] + Ji[r + 6] * Fi[2 + 3 * c];
    }

    if constexpr (sizeof(R) == 8) {
      // FP64: j^-T and detJ wait in LDS through the constitutive update (the
      // Q rows and this point's U slot are not written before it ends), so
      // that the law runs with 20 fewer live VGPRs
#pragma unroll
      for (int q = 0; q < 9; ++q) s_q(le)[q][g] = jci[q];
      s_u4(le)[g][0] = detJ;
    }
    // det F > 0 (false for NaN); det F <= 0 stops the run (NaN is left to
    // the residual check)
    const bool detf_pos = detF > R(0);
    if constexpr (SUR != 1) {
      if (detF <= R(0)) report_detf(A.health, e, (double)detF);
    }
    const int mat = A.mat_of_elem ? A.mat_of_elem[e] : 0;
    const double* pr = A.props + 18 * mat;
    const sma::Mat m = sma::make_mat_pre(pr, A.thresholds + 4 * mat, A.kg[2 * (int64_t)e],
                                         A.kg[2 * (int64_t)e + 1]);
    const int64_t gp = (int64_t)e * 8 + g;
    // a neo-Hookean element has no history: its values stay as they are in
    // both slots (equal from creation and every set), so they are neither
    // read nor written
    const bool neo = LAWS == 2 || (LAWS == 0 && A.law_of_elem && A.law_of_elem[e]);
    double xS0 = 0.0, FAS0 = 0.0, FSA0 = 0.0;
    if (!neo) {
      xS0 = A.sdv_n[0][gp];
      FAS0 = A.sdv_n[1][gp];
      FSA0 = A.sdv_n[2][gp];
    }

    R tau[6];
    R Fm[9];
    sma::ReturnMap rm;
    if (neo) {
      // neo-Hookean element: no history, no flags
      R jm1;
      sma::neo_hookean_tau<R>(pr[0], pr[1], H, tau, jm1);
      rm.xS = xS0;
      rm.F_AS = FAS0;
      rm.F_SA = FSA0;
      rm.flags = 0;
#pragma unroll
      for (int q = 0; q < 9; ++q) Fm[q] = H[q] + ((q % 4 == 0) ? R(1) : R(0));
      if constexpr (SUR == 1) A.sur_status[gp] = 0;
    } else if constexpr (sizeof(R) == 8) {
      // the Hencky tensor by series (theta, e' and e_n, no eigenvectors;
      // large stretches first halve log b by symmetric square roots)
      const sma::Hencky64 hk = sma::hencky_fp64(H);
      rm = sma::return_map(m, hk.theta, hk.e_n, xS0, FAS0, FSA0);
      sma::kirchhoff_hencky64(m, hk, rm.xS, tau);
      if (!hk.ok) rm.flags |= sma::FLAG_EIGEN_FAIL;
#pragma unroll
      for (int q = 0; q < 9; ++q) Fm[q] = H[q] + ((q % 4 == 0) ? 1.0 : 0.0);
    } else {
      float Hf[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) Hf[q] = H[q];
      // FP32: Hencky strain tensor by series (theta and |e'| are all the
      // return map and the stress need); the spectral form serves stretches
      // beyond the series' range
      const sma::Hencky32 hk = sma::hencky_fp32(Hf);
      if constexpr (SUR == 1) {
        double th = hk.theta, en = hk.e_n;
        sma::Spectral<float> sp{};
        if (!hk.ok) {
          sp = sma::spectral_fp32(Hf);
          th = sp.theta;
          en = sp.e_n;
        }
        bool kf_AS, kf_SA;
        double FAS, FSA;
        const bool flagged = sma::flags_only(m, th, en, xS0, FAS0, FSA0, kf_AS, kf_SA, FAS, FSA);
        // single-branch points whose features lie in the trained range
        // (|f| <= 8) go to the list
        float f[mlp::NIN];
        bool want = false;
        if (flagged && kf_AS != kf_SA) {
          // the features' principal deviatoric strains: in closed form from
          // the series' e' (the spectral path has them already)
          if (hk.ok) sma::principal_dev_fp32(hk.ed, sp.e[0], sp.e[1], sp.e[2]);
          const double E = A.E_of_elem ? A.E_of_elem[e] : pr[0];
          mlp::make_features(pr, E, A.T_feat, sp.e[0], sp.e[1], sp.e[2], th, xS0, FAS0, FSA0, FAS, f);
          want = true;
#pragma unroll
          for (int k = 0; k < mlp::NIN; ++k) want = want && fabsf(f[k]) <= 8.0f;
        }
        // one atomic per wave, on the counter of the wave's segment (wave
        // index mod LIST_SEGS; the wave covers elements 8 w .. 8 w + 7 of
        // the launch)
        const unsigned long long wm = __ballot(want);
        const int lane = threadIdx.x & 63;
        const int leader = wm ? __ffsll((long long)wm) - 1 : 0;
        const int seg = ((A.e_begin + blk * EPB) / 8 + (tid >> 6)) % mlp::LIST_SEGS;
        unsigned base = 0;
        if (wm && lane == leader)
          base = atomicAdd(A.sur_count + seg * mlp::LIST_STRIDE, (unsigned)__popcll(wm));
        base = __shfl(base, leader);
        if (want) {
          const int64_t slot = (int64_t)seg * A.sur_seg_cap + base +
                               (unsigned)__popcll(wm & ((1ull << lane) - 1ull));
          float4* dst = reinterpret_cast<float4*>(A.sur_feat + slot * mlp::NIN);
#pragma unroll
          for (int k = 0; k < mlp::NIN / 4; ++k)
            dst[k] = make_float4(f[4 * k], f[4 * k + 1], f[4 * k + 2], f[4 * k + 3]);
          A.sur_slot[gp] = (int32_t)slot;
        }
        A.sur_status[gp] = !flagged ? 0 : want ? 1 : 2;
      } else if (hk.ok) {
        if constexpr (SUR == 2) {
          rm = surrogate_or_exact(A, m, gp, hk.theta, hk.e_n, xS0, FAS0, FSA0);
        } else {
          rm = sma::return_map(m, hk.theta, hk.e_n, xS0, FAS0, FSA0);
        }
        sma::kirchhoff_dev(m, hk, rm.xS, tau);
      } else {
        sma::Spectral<float> sp = sma::spectral_fp32(Hf);
        if constexpr (SUR == 2) {
          rm = surrogate_or_exact(A, m, gp, sp.theta, sp.e_n, xS0, FAS0, FSA0);
        } else {
          rm = sma::return_map(m, sp.theta, sp.e_n, xS0, FAS0, FSA0);
        }
        sma::kirchhoff<float>(m, sp, rm.xS, tau);
        if (!sp.eig_ok) rm.flags |= sma::FLAG_EIGEN_FAIL;
      }
#pragma unroll
      for (int q = 0; q < 9; ++q) Fm[q] = H[q] + ((q % 4 == 0) ? R(1) : R(0));
    }
    if constexpr (SUR != 1) {
    if (!neo) {
      A.sdv_k[0][gp] = rm.xS;
      A.sdv_k[1][gp] = rm.F_AS;
      A.sdv_k[2][gp] = rm.F_SA;
    }
    fl = rm.flags;
    if (!detf_pos) fl |= sma::FLAG_DETF_NONPOS;

    // Q = detJ * tau * j^-T
    const R T[9] = {tau[0], tau[1], tau[2], tau[1], tau[3], tau[4], tau[2], tau[4], tau[5]};
    R jq[9], dJ;
    if constexpr (sizeof(R) == 8) {
#pragma unroll
      for (int q = 0; q < 9; ++q) jq[q] = s_q(le)[q][g];
      dJ = s_u4(le)[g][0];
    } else {
#pragma unroll
      for (int q = 0; q < 9; ++q) jq[q] = jci[q];
      dJ = detJ;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int k = 0; k < 3; ++k)
        s_q(le)[3 * i + k][g] =
            dJ * (T[i] * jq[k] + T[i + 3] * jq[k + 3] + T[i + 6] * jq[k + 6]);

    if constexpr (DAMP) {
      // Rayleigh damping at Gauss point g.  J and the velocity's natural
      // derivatives are formed again from the V(h) rows (they are not kept
      // through the constitutive update); the velocity interpolated at g is
      // gpvel = v_0 + 1/8 sum_h c^|h| s_g(h) V_v(h) (sld_elmpre.f90:110-114)
      const int gd = opaque(g);
      const R c1 = (R)(0.125 * HEX_GPOS), c2 = (R)(0.125 * HEX_GPOS * HEX_GPOS);
      const R c3 = (R)(0.125 * HEX_GPOS * HEX_GPOS * HEX_GPOS), k0 = (R)0.125;
      const R ax = sign_if(c1, ((gd >> 2) & 1) ^ 1), ay = sign_if(c1, ((gd >> 1) & 1) ^ 1),
              az = sign_if(c1, (gd & 1) ^ 1);
      const R byz = sign_if(c2, ((gd >> 1) ^ gd) & 1);
      const R bxz = sign_if(c2, ((gd >> 2) ^ gd) & 1);
      const R bxy = sign_if(c2, ((gd >> 2) ^ (gd >> 1)) & 1);
      const R bxyz = sign_if(c3, (((gd >> 2) ^ (gd >> 1) ^ gd) & 1) ^ 1);
      // weights of V(h): h = 0, z, y, yz, x, xz, xy, xyz (bits x = 4, y = 2, z = 1)
      const R wv[8] = {k0, az, ay, byz, ax, bxz, bxy, bxyz};
      const double alpha = A.damping[2 * mat], beta = A.damping[2 * mat + 1];
      const R arho = (R)(alpha * A.rho[mat]);
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        R gv = R(0);
#pragma unroll
        for (int h = 0; h < 8; ++h) gv += wv[h] * s_v(le)[h][6 + i];
        gv += (R)s_v0[le][i];
        s_q(le)[9 + i][g] = arho * dJ * gv;
      }
      R z[3] = {R(0), R(0), R(0)};
      if (beta > 0.0) {
        // the beta term as written (sld_elmmat.f90:174-194):
        // z_k = beta detJ sum_j J^-1_kj w_j, w_j = sum_k L_kj S_jk with the
        // velocity gradient L = G_v J^-1 (gprat, sld_elmpre.f90:150-157)
        R Jd[9], Gv[9];
        const R* V7 = s_v(le)[7];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const int hk = 4 >> k, h1 = k == 0 ? 2 : 4, h2 = k == 2 ? 2 : 1;
          const R w1 = k == 0 ? ay : ax, w2 = k == 2 ? ay : az, w3 = k == 0 ? byz : k == 1 ? bxz : bxy;
          const R* Vk = s_v(le)[hk];
          const R* V1 = s_v(le)[hk | h1];
          const R* V2 = s_v(le)[hk | h2];
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            Jd[i + 3 * k] = k0 * Vk[i] + w1 * V1[i] + w2 * V2[i] + w3 * V7[i];
            Gv[i + 3 * k] = k0 * Vk[6 + i] + w1 * V1[6 + i] + w2 * V2[6 + i] + w3 * V7[6 + i];
          }
        }
        R Jid[9];
        sma::inv3<R>(Jd, Jid);
        R L[9], Fi[9], S[9];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            L[r + 3 * c] = Gv[r] * Jid[3 * c] + Gv[r + 3] * Jid[1 + 3 * c] + Gv[r + 6] * Jid[2 + 3 * c];
            // F^-1 = J j^-1 (j = F J)
            Fi[r + 3 * c] = Jd[r] * jq[3 * c] + Jd[r + 3] * jq[1 + 3 * c] + Jd[r + 6] * jq[2 + 3 * c];
          }
        sma::pk2<R>(Fi, tau, S);
        R w[3];
#pragma unroll
        for (int j = 0; j < 3; ++j)
          w[j] = L[3 * j] * S[j] + L[1 + 3 * j] * S[j + 3] + L[2 + 3 * j] * S[j + 6];
        const R bd = (R)beta * dJ;
#pragma unroll
        for (int k = 0; k < 3; ++k)
          z[k] = bd * (Jid[k] * w[0] + Jid[k + 3] * w[1] + Jid[k + 6] * w[2]);
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) s_q(le)[12 + k][g] = z[k];
    }

    if constexpr (STORE) {
      R Fi[9], S[9];
      sma::inv3<R>(Fm, Fi);
      sma::pk2<R>(Fi, tau, S);
      const int64_t ngp = A.nelem * 8;
#pragma unroll
      for (int q = 0; q < 9; ++q) A.S_out[q * ngp + gp] = (double)S[q];
      // P_jk = sum_i F_ji S_ik (sld_builtin_materials.f90:355-371), column-major
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int j = 0; j < 3; ++j)
          A.P_out[(j + 3 * k) * ngp + gp] =
              (double)(Fm[j] * S[3 * k] + Fm[j + 3] * S[1 + 3 * k] + Fm[j + 6] * S[2 + 3 * k]);
      A.flags_out[gp] = fl;
    }
    }  // SUR != 1
  }
  if constexpr (SUR == 1) continue;   // uniform: the prepass has no scatter
  count_flags(A.counters, fl, valid);
  if constexpr (SUR == 2) {
    // the next step's candidates: elements with a transforming point now
    if (A.sur_cand_next) {
      const unsigned long long tm = __ballot(valid && (fl & 3u));
      if (valid && g == 0) A.sur_cand_next[e] = ((tm >> (8 * (le & 7))) & 0xffull) != 0ull;
    }
  }
  __syncthreads();
  // thread g: U of Q(j = g) (Q(8) is formed in phase 3 by every thread);
  // with damping threads 0-2 also form U of z_0..2 and threads 3-5 the
  // Walsh-Hadamard coefficients of q_0..2
  if (valid) {
    const int j = opaque(g);
    u4_of_q(s_q(le)[j], j % 3, s_u4(le)[j]);
    if constexpr (DAMP) {
      if (j < 3) {
        u4_of_q(s_q(le)[12 + j], j, s_u4(le)[8 + j]);
      } else if (j < 6) {
        R v[8];
#pragma unroll
        for (int b = 0; b < 8; ++b) v[b] = s_q(le)[9 + j - 3][b];
        wht_stage(v, 4);
        wht_stage(v, 2);
        wht_stage(v, 1);
#pragma unroll
        for (int h = 0; h < 8; ++h) s_wq[le][j - 3][h] = v[h];
      }
    }
  }
  __syncthreads();

  // -- phase 3: thread g now owns node a = g:
  // f_ai = sum_k s_ak/8 [U0 + c s_a1 U1 + c s_a2 U2 + c^2 s_a1 s_a2 U12](3i+k)
  if (valid) {
    const int a = opaque(g);
    const int sx = (0x66 >> a) & 1, sy = (0xCC >> a) & 1, sz = (0xF0 >> a) & 1;
    const R c0 = (R)0.125, c1 = (R)(0.125 * HEX_GPOS), c2 = (R)(0.125 * HEX_GPOS * HEX_GPOS);
    R u8[4];
    u4_of_q(s_q(le)[8], 2, u8);
    R f[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int sk = k == 0 ? sx : k == 1 ? sy : sz;
      const int s1 = k == 0 ? sy : sx, s2 = k == 2 ? sy : sz;
      // signs s_ak, s_ak s_a1, s_ak s_a2, s_ak s_a1 s_a2 as bit parities
      const R l0 = sign_if(c0, sk ^ 1);
      const R l1 = sign_if(c1, sk ^ s1);
      const R l2 = sign_if(c1, sk ^ s2);
      const R l3 = sign_if(c2, (sk ^ s1 ^ s2) ^ 1);
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const R* u = (3 * i + k == 8) ? u8 : s_u4(le)[3 * i + k];
        const R term = l0 * u[0] + l1 * u[1] + l2 * u[2] + l3 * u[3];
        f[i] = k == 0 ? term : f[i] + term;
      }
    }
    const R f0 = f[0], f1 = f[1], f2 = f[2];
    if constexpr (DAMP) {
      // damping force of corner a: alpha sum_g N_a(g) q(g) =
      // 1/8 sum_h c^|h| s_a(h) Q_q(h), less the beta sum
      // sum_g sum_k dN_a/dxi_k z_k (one value for the three components)
      R cb = R(0);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int sk = k == 0 ? sx : k == 1 ? sy : sz;
        const int s1 = k == 0 ? sy : sx, s2 = k == 2 ? sy : sz;
        const R l0 = sign_if(c0, sk ^ 1);
        const R l1 = sign_if(c1, sk ^ s1);
        const R l2 = sign_if(c1, sk ^ s2);
        const R l3 = sign_if(c2, (sk ^ s1 ^ s2) ^ 1);
        const R* u = s_u4(le)[8 + k];
        cb += l0 * u[0] + l1 * u[1] + l2 * u[2] + l3 * u[3];
      }
      const R c3 = (R)(0.125 * HEX_GPOS * HEX_GPOS * HEX_GPOS);
      const R px = sign_if(c1, sx ^ 1), py = sign_if(c1, sy ^ 1), pz = sign_if(c1, sz ^ 1);
      const R pyz = sign_if(c2, sy ^ sz), pxz = sign_if(c2, sx ^ sz), pxy = sign_if(c2, sx ^ sy);
      const R pxyz = sign_if(c3, (sx ^ sy ^ sz) ^ 1);
      const R wa[8] = {c0, pz, py, pyz, px, pxz, pxy, pxyz};
      R fd[3];
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        R r = R(0);
#pragma unroll
        for (int h = 0; h < 8; ++h) r += wa[h] * s_wq[le][i][h];
        fd[i] = r - cb;
      }
      R* ed = (R*)A.edamp + 3 * (int64_t)A.corner_slot[(int64_t)e * 8 + a];
      ed[0] = fd[0];
      ed[1] = fd[1];
      ed[2] = fd[2];
    }
    if constexpr (GATHER) {
      R* ef = (R*)A.efor + 3 * (int64_t)A.corner_slot[(int64_t)e * 8 + a];
      ef[0] = f0;
      ef[1] = f1;
      ef[2] = f2;
    } else {
      // coloured scatter: no other element of this colour touches this node
      R* fx = (R*)A.fint[0];
      R* fy = (R*)A.fint[1];
      R* fz = (R*)A.fint[2];
      fx[node] += f0;
      fy[node] += f1;
      fz[node] += f2;
    }
  }
  // the next tile's gather writes the LDS rows phase 3 has just read
  if (t_end - t_begin > 1) __syncthreads();
  }  // tile loop
}

// ----------------------------------------------------------------------------
// Setup: lumped mass (sld_elmmat.f90:199-209), one thread per element, per colour.
// ----------------------------------------------------------------------------
struct MassArgs {
  int32_t e_begin, e_end;       // colour range in elem_list
  const int32_t* elem_list;
  const int32_t* lnods;
  const double* coord[3];
  const double* rho;
  const int32_t* mat_of_elem;
  double* mass;
};

__global__ void sld_mass(MassArgs A) {
#pragma clang fp contract(off)
  const int i_el = A.e_begin + blockIdx.x * blockDim.x + threadIdx.x;
  if (i_el >= A.e_end) return;
  const int e = A.elem_list[i_el];
  double x[8][3];
  int nodes[8];
  for (int a = 0; a < 8; ++a) {
    nodes[a] = A.lnods[(int64_t)e * 8 + a];
    for (int d = 0; d < 3; ++d) x[a][d] = A.coord[d][nodes[a]];
  }
  const double rho = A.rho[A.mat_of_elem ? A.mat_of_elem[e] : 0];
  double elmuu[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int g = 0; g < 8; ++g) {
    double J[9] = {0};
    for (int a = 0; a < 8; ++a)
      for (int k = 0; k < 3; ++k) {
        double D = hex_deriv<double>(a, g, k);
        for (int i = 0; i < 3; ++i) J[i + 3 * k] = J[i + 3 * k] + x[a][i] * D;
      }
    // jacobi.f90 ndime==3 determinant by first-row cofactors
    double t1 = J[4] * J[8] - J[5] * J[7];
    double t2 = -J[1] * J[8] + J[2] * J[7];
    double t3 = J[1] * J[5] - J[2] * J[4];
    double vol = 1.0 * (J[0] * t1 + J[3] * t2 + J[6] * t3);
    for (int a = 0; a < 8; ++a)
      for (int b = 0; b < 8; ++b)
        elmuu[a] = elmuu[a] + rho * vol * hex_shape<double>(a, g) * hex_shape<double>(b, g);
  }
  for (int a = 0; a < 8; ++a) A.mass[nodes[a]] += elmuu[a];
}

// Setup: bulk and shear modulus per element (the expressions of make_mat).
__global__ void sld_elem_moduli(int64_t ne, const double* __restrict__ props,
                                const int32_t* __restrict__ mat_of_elem,
                                const double* __restrict__ E_of_elem, double* __restrict__ kg) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const double* pr = props + 18 * (mat_of_elem ? mat_of_elem[e] : 0);
  sma::elastic_moduli(pr, E_of_elem ? E_of_elem[e] : pr[0], kg[2 * e], kg[2 * e + 1]);
}

__global__ void sld_reciprocal(const double* __restrict__ m, double* __restrict__ rm, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) rm[i] = 1.0 / m[i];
}

// ----------------------------------------------------------------------------
// Point-wise law (drop-in for eval_sld_stress_model_160 with flagt = 0)
// ----------------------------------------------------------------------------
struct LawArgs {
  int64_t n;
  int32_t nmat;
  const double* props;
  const int32_t* mat_of_gp;
  const double* E_of_gp;
  const double* F;        // [9][n]
  const double* sdv_old;  // [3][n]
  double* sdv_new;        // [3][n]
  double* S;              // [9][n]
  uint32_t* flags;        // [n] or null
};

template <bool FP32_KIN>
__global__ void __launch_bounds__(256) sma160_points(LawArgs A) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  int mat = A.mat_of_gp ? A.mat_of_gp[i] : 0;
  uint32_t bad = 0;
  if (mat < 0 || mat >= A.nmat) {  // device arrays cannot be checked on the host
    bad = sma::FLAG_BAD_INPUT;
    mat = 0;
  }
  const double* pr = A.props + 18 * mat;
  const double E = A.E_of_gp ? A.E_of_gp[i] : pr[0];
  const sma::Mat m = sma::make_mat(pr, E);
  double F[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) F[q] = A.F[q * A.n + i];
  const double xS0 = A.sdv_old[i], FAS0 = A.sdv_old[A.n + i], FSA0 = A.sdv_old[2 * A.n + i];
  sma::ReturnMap rm;
  double S[9];
  if constexpr (!FP32_KIN) {
    sma::Spectral<double> sp = sma::spectral_fp64(F);
    rm = sma::return_map(m, sp.theta, sp.e_n, xS0, FAS0, FSA0);
    if (!sp.eig_ok) rm.flags |= sma::FLAG_EIGEN_FAIL;
    double tau[6], Fi[9];
    sma::kirchhoff<double>(m, sp, rm.xS, tau);
    double det = sma::inv3<double>(F, Fi);
    if (!(det > 0.0)) rm.flags |= sma::FLAG_DETF_NONPOS;
    sma::pk2<double>(Fi, tau, S);
  } else {
    float H[9], Ff[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) {
      H[q] = (float)(F[q] - ((q % 4 == 0) ? 1.0 : 0.0));
      Ff[q] = (float)F[q];
    }
    sma::Spectral<float> sp = sma::spectral_fp32(H);
    rm = sma::return_map(m, sp.theta, sp.e_n, xS0, FAS0, FSA0);
    if (!sp.eig_ok) rm.flags |= sma::FLAG_EIGEN_FAIL;
    float tau[6], Fi[9], Sf[9];
    sma::kirchhoff<float>(m, sp, rm.xS, tau);
    float det = sma::inv3<float>(Ff, Fi);
    if (!(det > 0.f)) rm.flags |= sma::FLAG_DETF_NONPOS;
    sma::pk2<float>(Fi, tau, Sf);
#pragma unroll
    for (int q = 0; q < 9; ++q) S[q] = Sf[q];
  }
  A.sdv_new[i] = rm.xS;
  A.sdv_new[A.n + i] = rm.F_AS;
  A.sdv_new[2 * A.n + i] = rm.F_SA;
#pragma unroll
  for (int q = 0; q < 9; ++q) A.S[q * A.n + i] = S[q];
  if (A.flags) A.flags[i] = rm.flags | bad;
}

}  // namespace sld
