/*
 * TEST INFRASTRUCTURE ONLY -- the CPU oracle.
 *
 * A plain-C FP64 restatement of the reference's explicit solid-mechanics
 * step with the superelastic SMA law (Alya solidz, stress model 160).  It is
 * the checker for the HIP path and the CPU baseline timed beside it; only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * Nothing in the product library links or calls it.
 *
 * Functions and what pins each (see tests/test_oracle_*.py):
 *   orc_sm160_eval        sld_stress_model_160.f90:36-195 + :212-614,
 *                         dsyevq3.f90, dsytrd3.f90, eigsrt.f90.
 *                         Pinned by golden vectors written by the reference's
 *                         own compiled law (tests/golden/), by the Hencky
 *                         closed form below threshold and at saturation,
 *                         frame indifference and the Newton-fail revert.
 *   orc_hex8_tables       mod_lagrange.f90:935-1000, mod_quadrature_gauss.f90:
 *                         2320-2321,2372-2385.  Pinned by golden vectors from
 *                         the reference's shape3/gauss_qua and partition of unity.
 *   orc_hex8_kinematics   jacobi.f90 (ndime==3), elmca2.f90:105-111,
 *                         sld_elmpre.f90:149-162,219-226, invmtx.f90.
 *                         Pinned by golden vectors from the reference's jacobi/
 *                         invmtx and by exact affine fields.
 *   orc_sld_*             element RHS, lumped mass, integrators, load tables:
 *                         sld_elmmat.f90:115-209, sld_builtin_materials.f90:
 *                         355-371, mod_sld_solution_methods.f90:101-223,
 *                         385-574, solric.f90:294-330, sld_funbou.f90:95-136.
 *                         Pinned by patch tests (rigid motion, constant
 *                         stress), mass = rho*V, virtual work, the closed-form
 *                         harmonic oscillator and the smooth-step formula.
 */
#ifndef ALYA_SMA_ORACLE_H
#define ALYA_SMA_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Per-point status bits; identical values to include/alya_sma.h. */
#define ORC_FLAG_KF_AS        0x01u
#define ORC_FLAG_KF_SA        0x02u
#define ORC_FLAG_NEWTON_FAIL  0x04u
#define ORC_FLAG_EIGEN_FAIL   0x08u
#define ORC_FLAG_DETF_NONPOS  0x10u
#define ORC_FLAG_ITER_SHIFT   8        /* Newton iterations in bits 8..15 */

/* Point-wise SMA law on SoA arrays (component k of point i at [k*n + i]).
 * props: [n_mat][18]; mat_of_gp may be NULL (material 0 everywhere).
 * E_of_gp may be NULL; when given it overrides props[0] per point.
 * F and S are 3x3 column-major per point (k = row + 3*col).
 * sdv_old/sdv_new: (xS, F_AS, F_SA).  Returns 0. */
int orc_sm160_eval(int64_t n, const double *props, const int32_t *mat_of_gp,
                   const double *E_of_gp, const double *F, const double *T,
                   const double *sdv_old, double *sdv_new, double *S,
                   uint32_t *flags);

/* Single point, AoS, also returns the Kirchhoff stress tau (col-major). */
uint32_t orc_sm160_point(const double props[18], const double F[9],
                         const double sdv_old[3], double sdv_new[3],
                         double S[9], double tau[9]);

/* Kopp's QL eigen-solver restated (A read-only, col-major); returns 0 on
 * convergence, 1 on the 30-iteration bail-out.  eigsrt descending. */
int orc_dsyevq3(const double *A, double *Q, double w[3]);
void orc_eigsrt(double d[3], double *V);

/* HEX08 tables at the 2x2x2 Gauss points: posgp[8][3], weigp[8],
 * shapf[8 gp][8 node], deriv[8 gp][8 node][3]. */
void orc_hex8_tables(double *posgp, double *weigp, double *shapf, double *deriv);

/* Per-element kinematics at the 8 GPs.  elcod/eldis: [8 node][3].
 * Outputs gpcar[8 gp][8 node][3], gpvol[8], F[8 gp][9 col-major], detF[8]. */
void orc_hex8_kinematics(const double *elcod, const double *eldis,
                         double *gpcar, double *gpvol, double *F, double *detF);

/* ------------------------------------------------------------------ */
/* Whole explicit step                                                  */
/* ------------------------------------------------------------------ */
typedef struct {
    int64_t npoin, nelem;
    const int32_t *lnods;      /* [nelem][8], 0-based, reference node order */
    const double  *coord;      /* [npoin][3] reference coordinates          */
    int32_t nmat;
    const double  *props;      /* [nmat][18] SMA parameters                 */
    const double  *rho;        /* [nmat] density                            */
    const int32_t *mat_of_elem;/* [nelem] or NULL                           */
    const double  *E_of_elem;  /* [nelem] Young's modulus override or NULL  */
    const uint8_t *fixmask;    /* [npoin] bit d set = dof d prescribed       */
    const int32_t *func_of_node; /* [npoin] Dirichlet function id or -1     */
    int32_t nfunc;
    const int32_t *func_npts;  /* [nfunc]                                   */
    const int32_t *func_smooth;/* [nfunc] 1 = DISCRETE,SMOOTH, 0 = linear   */
    const double  *func_table; /* [nfunc][maxpts][4] (v_x, v_y, v_z, t)      */
    int32_t func_maxpts;
    const double  *fext0;      /* [npoin][3] nodal force pattern or NULL     */
    int32_t fext_func;         /* load function scaling fext0 (component x)  */
    const int32_t *elem_order; /* [nelem] processing order or NULL           */
    const int32_t *color_ptr;  /* [ncolor+1] ranges into elem_order or NULL  */
    int32_t ncolor;
    const int32_t *law_of_mat; /* [nmat] 160 SMA, 102 neo-Hookean; NULL = 160 */
    /* Neumann loads on QUA04 boundary faces (sld_bouope.f90) */
    int64_t nface;
    const int32_t *lnodb;      /* [nface][4] 0-based                         */
    const int32_t *face_code;  /* [nface] 3 pressure along the normal, 6 traction */
    const double  *face_value; /* [nface][3] constant values (bvnat)         */
    const int32_t *face_func;  /* [nface] function id or -1; NULL = none     */
    int32_t face_transient;    /* 1: value = (function column x, 0, 0) or 0  */
    int32_t face_follower;     /* 1: integrate over X + u^{n+1}              */
    /* gravity acceleration g = grnor_sld * gravi_sld (sld_reaphy.f90:209-223);
     * its element load sum_g N_a rho g w detJ (sld_elmmat.f90:152-157) is the
     * lumped mass times g */
    int32_t has_gravity;
    double gravity[3];
    /* Rayleigh damping dampi_sld(1:2, imate) = (alpha, beta) per material
     * (RAYLEIGH_DAMPING: RAYALPHA, RAYBETA, sld_reaphy.f90:1046-1049), or
     * NULL; terms of sld_elmmat.f90:160-163 and :174-194 */
    const double *damping;     /* [nmat][2] */
    /* body force bodyf_sld(t^{n+1}) of load function body_func (BODYF,
     * sld_reabcs.f90:523; evaluated by sld_funbou.f90 itask 3 in
     * sld_updbcs.f90:96-99), -1 = none; element load of sld_elmmat.f90:
     * 156-157, the lumped mass times the body force */
    int32_t body_func;
} orc_model;

/* Nodal forces of the boundary faces at time t over the geometry
 * coord (+ u when follower): ff[npoin][3], overwritten. */
void orc_face_loads(const orc_model *m, const double *u, double t, double *ff);

/* Neo-Hookean (Ansys) PK2 at one point: props[0] = mu, props[1] = K,
 * F column-major, J = det F (sld_stress_model_102.f90:60-119). */
void orc_nh102_point(const double props[18], const double F[9], double J, double S[9]);

typedef struct {
    double *u, *v, *a;         /* [npoin][3] (ITER_K after a step)           */
    double *fint;              /* [npoin][3] assembled internal force        */
    double *rhs;               /* [npoin][3] assembled RHS = fext - fint     */
    double *mass;              /* [npoin] lumped mass                        */
    double *sdv;               /* [2][nelem][8][3]: [0]=ITER_K, [1]=TIME_N   */
    double *S;                 /* [nelem][8][9] PK2 of the last step or NULL */
    uint32_t *flags;           /* [nelem][8] or NULL                         */
    double t;                  /* current time                               */
    int64_t nsteps;
    /* energy balance (mod_sld_energy.f90:71-162); accumulated when
     * fint_prev != NULL ([npoin][3], f_int of the previous step) */
    double *fint_prev;
    double *fext_prev;         /* [npoin][3] f_ext of the previous step      */
    double allie, allwk, allke, etota;
    /* run stop (sld_detjac / sld_cvgunk runend): 0, or 1 = det F <= 0 at a
     * Gauss point (the step ends after the element loop), 2 = a NaN or
     * infinite predicted displacement (the step ends after the predictor)
     * or residual (after the corrector) */
    int32_t halt;
    int64_t halt_elem;         /* element of the smallest det F (first in loop order) */
    double halt_detf;
    int64_t halt_node;         /* lowest node with a non-finite value */
} orc_state;

int orc_sld_mass(const orc_model *m, orc_state *s);

/* Phases of the step, for the FLOP-counting build (oracle/flopcount.cpp,
 * which defines ORC_PHASE); a no-op in the oracle library. */
#ifndef ORC_PHASE
#define ORC_PHASE(k) ((void)0)
#endif
#define ORC_PH_NODAL 0   /* predictor, loads, solve, corrector */
#define ORC_PH_KIN   1   /* element gather, Jacobian, gpcar, F, det F */
#define ORC_PH_LAW   2   /* constitutive law at a Gauss point */
#define ORC_PH_FORCE 3   /* P = F S, element internal force, assembly */

/* Planted mistakes, for tests that show a check catches them: 0 (the
 * default) runs the restatement as written; the tests set another value,
 * check that their closed form fails, and set it back.  Never set outside
 * such a test. */
#define ORC_MUT_NONE        0
#define ORC_MUT_DT_YOUNG    1   /* sound speed sqrt(E/rho) for sqrt((lam+2mu)/rho) */
#define ORC_MUT_DT_HMAX     2   /* hleng(1), the largest length, for hleng(ndime) */
#define ORC_MUT_DT_NH_BULK  3   /* law 102: sqrt(K/rho), the 4 mu/3 dropped */
#define ORC_MUT_TW_HALF     4   /* TW predictor with phi dt^2 a / 2 */
#define ORC_MUT_TW_FIXVA    5   /* TW corrector keeps v, a on prescribed dofs */
#define ORC_MUT_TW_VNEW     6   /* TW predictor u with v^{n+1} for v^n */
#define ORC_MUT_DAMP_VHALF  7   /* CD damping with v^{n+1/2} for veloc_sld(ITER_K) = v^n */
#define ORC_MUT_DAMP_LUMPED 8   /* alpha term with the lumped mass (node velocity) */
#define ORC_MUT_BETA_T      9   /* beta term with the velocity gradient transposed */
#define ORC_MUT_BODY_LE     10  /* body-force table with t <= t2 for t < t2 */
extern int orc_mutation;

double orc_sld_critical_dt(const orc_model *m);
/* OpenMP threads of the element loop (no-op without OpenMP). */
void orc_set_threads(int n);
/* One explicit step. scheme 0 = central difference, 1 = Tchamwa-Wielgosz.
 * Returns s->halt (0 = completed). */
int orc_sld_step(const orc_model *m, orc_state *s, double dt, int scheme, double phi);
/* Evaluate load function f at time t (sld_funbou.f90 itask 2, component c). */
double orc_funbou(const orc_model *m, int32_t f, int c, double t);
/* Body-force value of load function f at time t (sld_funbou.f90 itask 3,
 * component c): linear between table rows, zero outside the table. */
double orc_funbou_body(const orc_model *m, int32_t f, int c, double t);

#ifdef __cplusplus
}
#endif
#endif
