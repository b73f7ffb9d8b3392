/*
 * TEST INFRASTRUCTURE ONLY -- algorithmic FLOP count of the explicit step,
 * by instrumenting the CPU oracle: its C sources are compiled here as C++
 * with `double` standing for F64 (oracle/flop_f64.hpp), which counts every
 * +, -, *, / (and sqrt, log, pow apart) in the reference's operation order.
 *
 *   make -C oracle flopcount && oracle/_build/flopcount
 *
 * Counts, per Gauss point of a HEX08 element unless noted:
 *   kinematics   gather, Jacobian and its inverse, gpcar, F, det F
 *                (elmca2/jacobi, sld_elmpre), per Gauss point
 *   law          sm160 at a point (sld_stress_model_160.f90:99-192):
 *                elastic (no branch), austenite->martensite loading (one
 *                Newton update), martensite->austenite unloading; and the
 *                neo-Hookean law 102 (sld_stress_model_102.f90:100-119)
 *   force        P = F S and the element internal force with its assembly
 *                (sld_builtin_materials.f90:355-371, sld_elmmat.f90:137-150,
 *                mod_matrix.f90 assembly), per Gauss point
 *   nodal        predictor, solve, corrector, per node
 * The element is a distorted hex under a homogeneous stretch, so all eight
 * Gauss points take the same branch.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <cmath>

#include "flop_f64.hpp"

namespace flop {
long long count[NPHASE][NKIND];
int phase = 0;
}  // namespace flop

#define ORC_PHASE(k) (flop::phase = (k))
#define double F64
#include "sma160.c"
#include "hex8.c"
#include "nh102.c"
#include "sld_step.c"
#undef double

static void reset() { memset(flop::count, 0, sizeof(flop::count)); }

static long long flops(int ph) {
  return flop::count[ph][flop::ADD] + flop::count[ph][flop::MUL] + flop::count[ph][flop::DIV];
}

static void report(const char* name, int ph, double per) {
  const long long* c = flop::count[ph];
  printf("%-26s %8.1f  add %.1f mul %.1f div %.1f sqrt %.1f log %.1f pow %.1f\n", name,
         flops(ph) / per, c[flop::ADD] / per, c[flop::MUL] / per, c[flop::DIV] / per,
         c[flop::SQRT] / per, c[flop::LOG] / per, c[flop::POW] / per);
}

int main() {
  // one distorted HEX08 element, all nodes free, Nitinol (E 60 GPa)
  const double X[8][3] = {{0.0, 0.0, 0.0},  {1.1, 0.05, 0.0}, {1.0, 0.9, 0.1},  {-0.1, 1.0, 0.0},
                          {0.05, 0.0, 1.0}, {1.0, 0.1, 1.05}, {1.05, 1.0, 0.95}, {0.0, 0.95, 1.0}};
  F64 coord[24], props[18] = {60000.0, 0.33, 0.05, 400.0, 500.0, 0, 0, 0, 0,
                              250.0, 150.0, 0, 0, 0, 0, 1.0, 1.0, 0};
  for (int a = 0; a < 8; a++)
    for (int d = 0; d < 3; d++) coord[3 * a + d] = X[a][d];
  int32_t lnods[8] = {0, 1, 2, 3, 4, 5, 6, 7};
  F64 rho[1] = {6.45e-9};
  uint8_t fix[8] = {0};
  F64 ftab[4] = {0, 0, 0, 0};
  int32_t fnpts[1] = {1}, fsm[1] = {0};
  orc_model m;
  memset(&m, 0, sizeof(m));
  m.npoin = 8;
  m.nelem = 1;
  m.lnods = lnods;
  m.coord = coord;
  m.nmat = 1;
  m.props = props;
  m.rho = rho;
  m.fixmask = fix;
  m.nfunc = 1;
  m.func_npts = fnpts;
  m.func_smooth = fsm;
  m.func_table = ftab;
  m.func_maxpts = 1;
  m.fext_func = -1;
  m.body_func = -1;
  F64 u[24], v[24], a[24], fint[24], rhs[24], mass[8], sdv[2 * 24], S[72];
  uint32_t flags[8];
  orc_state s;
  memset(&s, 0, sizeof(s));
  s.u = u; s.v = v; s.a = a; s.fint = fint; s.rhs = rhs; s.mass = mass;
  s.sdv = sdv; s.S = S; s.flags = flags;
  memset(v, 0, sizeof(v));
  memset(a, 0, sizeof(a));
  memset(sdv, 0, sizeof(sdv));
  orc_sld_mass(&m, &s);
  (void)shapf_at(0, 0);   // the shape-function table is set-up work, not counted

  // a homogeneous stretch u = (F0 - I) X: every Gauss point sees F0
  auto stretch = [&](double ex) {
    memset(v, 0, sizeof(v));
    memset(a, 0, sizeof(a));
    const double F0[3] = {1.0 + ex, 1.0 - 0.3 * ex, 1.0 - 0.3 * ex};
    for (int p = 0; p < 8; p++)
      for (int d = 0; d < 3; d++) u[3 * p + d] = (F0[d] - 1.0) * X[p][d];
  };
  const double dt = 1e-9;

  printf("# algorithmic FLOPs of the oracle (reference operation order); "
         "add counts + and -; FLOPs = add + mul + div\n");
  // elastic step
  stretch(0.002);
  reset();
  orc_sld_step(&m, &s, dt, 0, 1.0);
  if (flags[0] & 3u) { fprintf(stderr, "elastic case transformed\n"); return 1; }
  report("kinematics_per_gp", ORC_PH_KIN, 8.0);
  report("law_elastic_per_gp", ORC_PH_LAW, 8.0);
  report("force_per_gp", ORC_PH_FORCE, 8.0);
  report("nodal_per_node", ORC_PH_NODAL, 8.0);

  // loading branch (austenite -> martensite) from a fresh history
  memset(sdv, 0, sizeof(sdv));
  s.halt = 0;
  stretch(0.03);
  reset();
  orc_sld_step(&m, &s, dt, 0, 1.0);
  if ((flags[0] & 3u) != 1u) { fprintf(stderr, "loading case: flags %u\n", flags[0]); return 1; }
  printf("# loading: Newton iterations %u\n", (flags[0] >> 8) & 0xffu);
  report("law_AS_per_gp", ORC_PH_LAW, 8.0);

  // unloading branch (martensite -> austenite): the loaded history (TIME_N
  // after the step), less strain
  stretch(0.012);
  reset();
  orc_sld_step(&m, &s, dt, 0, 1.0);
  if ((flags[0] & 3u) != 2u) { fprintf(stderr, "unloading case: flags %u\n", flags[0]); return 1; }
  printf("# unloading: Newton iterations %u\n", (flags[0] >> 8) & 0xffu);
  report("law_SA_per_gp", ORC_PH_LAW, 8.0);

  // neo-Hookean law 102 (mu 0.03, K 1 MPa)
  F64 nh[18] = {0.03, 1.0};
  int32_t law[1] = {102};
  m.props = nh;
  m.law_of_mat = law;
  stretch(0.01);
  reset();
  orc_sld_step(&m, &s, dt, 0, 1.0);
  report("law_nh102_per_gp", ORC_PH_LAW, 8.0);
  return 0;
}
