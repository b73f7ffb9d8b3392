/*
 * alya_sma.h -- C ABI of the MI355X explicit solid-mechanics step with the
 * superelastic SMA law (Alya solidz, stress model 160 "NITINOL").
 *
 * The reference exposes Fortran module procedures only; every entry point
 * below names the reference procedure(s) it replaces.  Conventions:
 *   - plain pointers and sizes, no C++ or torch types;
 *   - int return status: 0 = ok, < 0 = error (text from sma_last_error());
 *   - 3x3 tensors are column-major per point (k = row + 3*col), as Fortran
 *     stores gpgdi(3,3,pgaus) and gpstr(3,3,pgaus);
 *   - point arrays are SoA: component k of point i at [k*n + i];
 *   - history per Gauss point is (xS, F_AS, F_SA), the reference's
 *     svegm_sld(ielem)%a(1:3, igaus, :).
 */
#ifndef ALYA_SMA_H
#define ALYA_SMA_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Per-Gauss-point status bits. */
#define SMA_FLAG_KF_AS          0x01u  /* austenite->martensite branch (sld_stress_model_160.f90:351) */
#define SMA_FLAG_KF_SA          0x02u  /* martensite->austenite branch (:368) */
#define SMA_FLAG_NEWTON_FAIL    0x04u  /* 65-iteration cap, history reverted (:459-465) */
#define SMA_FLAG_EIGEN_FAIL     0x08u  /* DSYEVQ3 30-iteration bail-out (dsyevq3.f90:82-84) */
#define SMA_FLAG_DETF_NONPOS    0x10u  /* det F <= 0 (sld_elmpre.f90:229-240) */
#define SMA_FLAG_SURR_USED      0x20u  /* surrogate answered this point */
#define SMA_FLAG_SURR_FALLBACK  0x40u  /* surrogate check failed, exact law used */
#define SMA_FLAG_BAD_INPUT      0x80u  /* material index out of range (material 0 used) */
#define SMA_FLAG_ITER_SHIFT     8      /* Newton iterations in bits 8..15 */

/* Status codes. */
#define SMA_OK          0
#define SMA_ERR_ARG    -1
#define SMA_ERR_HIP    -2
#define SMA_ERR_STATE  -3

/* Text of the last error on the calling thread ("" if none). */
const char *sma_last_error(void);

/* ------------------------------------------------------------------------
 * Point-wise constitutive law.
 * Replaces eval_sld_stress_model_160(pgaus, props, gpgdi, gptmp, gpsdv,
 *   gpstr, gpdds, flagt=0)  (sld_stress_model_160.f90:36-48, called from
 *   sld_builtin_materials.f90:327-335).
 * props [n_mat][18]: the reference's parco_sld(1:18, pmate) layout.
 * mat_of_gp [n_gp] may be NULL (material 0); E_of_gp may be NULL (else it
 *   overrides props[0] per point).  T is accepted and ignored, as the
 *   reference hard-codes gp_tmp = 0 (:105).
 * F [9][n_gp] in; sdv_old [3][n_gp] = TIME_N in; sdv_new [3][n_gp] = ITER_K
 *   out; S [9][n_gp] second Piola-Kirchhoff out; flags [n_gp] out (may be NULL).
 * precision: 0 = FP64, 1 = FP32 spectral/stress with FP64 return map.
 * Host pointers; synchronous.  Each host thread keeps one device block and
 * one pinned host block, grown to the largest call and held to the process's
 * end, so that repeated small calls (an element's points) allocate nothing.
 * ------------------------------------------------------------------------ */
int sma160_eval(int64_t n_gp, int32_t n_mat, const double *props,
                const int32_t *mat_of_gp, const double *E_of_gp,
                const double *F, const double *T, const double *sdv_old,
                double *sdv_new, double *S, uint32_t *flags, int32_t precision);

/* Same as sma160_eval with every array (props included) a device pointer,
 * enqueued on `stream` (a hipStream_t, NULL = default stream). */
int sma160_eval_device(int64_t n_gp, int32_t n_mat, const double *props,
                       const int32_t *mat_of_gp, const double *E_of_gp,
                       const double *F, const double *T, const double *sdv_old,
                       double *sdv_new, double *S, uint32_t *flags,
                       int32_t precision, void *stream);

/* Same arguments and results as sma160_eval (host pointers, synchronous,
 * null stream), with no device allocation: the arrays are staged through a
 * fixed array in the library's device memory (3.4 MB on each GPU it runs
 * on, held until the process ends) instead of buffers allocated and freed
 * per call.  No counterpart in the reference: it serves callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  Rejects n_gp > SMA160_NOALLOC_MAX_GP
 * and n_mat > SMA160_NOALLOC_MAX_MAT with SMA_ERR_ARG.  Calls that overlap
 * on one GPU (from two host threads) race on the shared staging array. */
#define SMA160_NOALLOC_MAX_GP  16384
#define SMA160_NOALLOC_MAX_MAT 64
int sma160_eval_noalloc(int64_t n_gp, int32_t n_mat, const double *props,
                        const int32_t *mat_of_gp, const double *E_of_gp,
                        const double *F, const double *T, const double *sdv_old,
                        double *sdv_new, double *S, uint32_t *flags, int32_t precision);

/* ------------------------------------------------------------------------
 * Explicit step on the device.
 * Replaces the per-step call chain sld_explicit / sld_explicit_tw
 *   (mod_sld_solution_methods.f90:101-223, :385-574) -> sld_matrix
 *   (sld_matrix.f90:29-327) -> sld_elmope (sld_elmope.f90:53-99) ->
 *   sld_element_operations (sld_element_operations.f90:24-407) ->
 *   solver_solve/solric local mass (solric.f90:294-330), plus
 *   sld_updbcs(ITASK_BEGSTE) (sld_updbcs.f90:79-114) and
 *   sld_updunk(ITASK_ENDSTE) (sld_updunk.f90:149-186).
 * ------------------------------------------------------------------------ */
typedef struct sld_dev sld_dev;

typedef struct {
  int64_t npoin, nelem;
  const int32_t *lnods;        /* [nelem][8] 0-based, HEX08 node order of mod_lagrange.f90:946-953 */
  const double *coord;         /* [npoin][3] reference coordinates (coord(ndime,npoin)) */
} sld_mesh_desc;

typedef struct {
  int32_t nmat;
  const double *props;         /* [nmat][18] parco_sld(1:18, imate) */
  const double *rho;           /* [nmat] densi_sld(1, imate) */
  const int32_t *mat_of_elem;  /* [nelem] lmate_sld, or NULL */
  const double *E_of_elem;     /* [nelem] per-element Young's modulus (SMA), or NULL */
  const int32_t *law;          /* [nmat] lawst_sld: SLD_LAW_SMA160 or SLD_LAW_NEOHOOKEAN102;
                                  NULL = all SMA.  Neo-Hookean rows: props[0] = mu,
                                  props[1] = K (NEOHO ANSYS, sld_stress_model_102.f90:60-61) */
} sld_material_desc;

#define SLD_LAW_SMA160         160
#define SLD_LAW_NEOHOOKEAN102  102

typedef struct {
  const uint8_t *fixmask;      /* [npoin] bit d = kfl_fixno_sld(d+1, ipoin) == 1 */
  const int32_t *func_of_node; /* [npoin] load function index or -1 (kfl_funno_sld) */
  int32_t nfunc, func_maxpts;
  const int32_t *func_npts;    /* [nfunc] mtloa_sld */
  const int32_t *func_smooth;  /* [nfunc] 1 = DISCRETE,SMOOTH (kfl_funty_sld(2,:)) */
  const double *func_table;    /* [nfunc][func_maxpts][4] rows (v_x, v_y, v_z, t) = tload_sld */
  const double *fext0;         /* [npoin][3] nodal force pattern, or NULL */
  int32_t fext_func;           /* function scaling fext0 (component x), -1 = none */
} sld_bc_desc;

/* Builds the device state: validates indices, colours the elements
 * (greedy, deterministic), uploads SoA arrays, assembles the lumped mass.
 * Allocates on the current HIP device. */
int sld_dev_create(const sld_mesh_desc *mesh, const sld_material_desc *mat,
                   const sld_bc_desc *bc, sld_dev **out);

/* Precision mode: 0 = FP64 (plain), 1 = FP32 kinematics/assembly with FP64
 * return map and FP64 u/v/a accumulators, 2 = mode 1 + MLP surrogate
 * (requires sld_dev_set_surrogate). */
#define SLD_MODE_FP64       0
#define SLD_MODE_FP32       1
#define SLD_MODE_SURROGATE  2
int sld_dev_set_mode(sld_dev *dev, int32_t mode);

/* Store PK2 stress and status flags per Gauss point every step (for
 * checking); off by default. */
int sld_dev_set_store(sld_dev *dev, int32_t store);

/* Advance nsteps explicit steps of size dt.  scheme 0 = central difference
 * (sld_explicit), 1 = Tchamwa-Wielgosz with parameter phi (sld_explicit_tw).
 * Asynchronous on the handle's stream; steps are replayed from a hipGraph.
 * Failure stops the run as the reference's runend does: det F <= 0 at a
 * Gauss point (sld_elmope.f90:106 -> sld_detjac.f90:93) or a NaN/infinite
 * displacement or residual (sld_cvgunk.f90:195-199).  The kernel that finds
 * it halts the device (later steps do nothing, the state stays as the
 * failure left it) and the next synchronising call -- sld_dev_sync,
 * sld_dev_time_kernels, sld_dev_autotune -- returns SMA_ERR_STATE naming the
 * element (most negative det F) or node; from then on sld_dev_step returns
 * SMA_ERR_STATE at once.  sld_dev_get still reads the fields (and
 * SLD_FIELD_HEALTH); sld_dev_set of the state clears the halt. */
int sld_dev_step(sld_dev *dev, double dt, int32_t scheme, double phi, int64_t nsteps);

/* Builds (captures and instantiates) the graphs that sld_dev_step with the
 * same arguments would replay from the current state and settings, and runs
 * no step: the state is unchanged.  No counterpart in the reference (whose
 * steps are plain calls); for callers that time or latency-bound a run of
 * steps and want its one-time graph set-up out of it.  sld_dev_step builds
 * whatever is missing itself, so the call is optional. */
int sld_dev_prepare(sld_dev *dev, double dt, int32_t scheme, double phi, int64_t nsteps);

/* Runs nsteps steps eagerly with HIP events around the predictor, the
 * element phase and the corrector on the handle's stream (the reference
 * brackets the same phases with its times(3..7) timers,
 * sld_element_operations.f90:125-256).  With the gather assembly the
 * element launch is repeated 8 times back to back inside the bracket (same
 * inputs, same outputs; flag counts of the repeats are discarded).  out[8]:
 * ms per launch of the predictor, the element kernel (bracket / launches
 * in it) and the corrector, element launches per step, kernel ms per step,
 * eager stream ms per step (repeats included), and in surrogate mode ms per
 * launch of the prepass and of the inference (0 in the other modes).  The
 * steps are enqueued back
 * to back; with nsteps > 1 the first bracket (opened on an idle stream) is
 * left out of the averages. */
int sld_dev_time_kernels(sld_dev *dev, double dt, int32_t scheme, double phi,
                         int32_t nsteps, double *out);

/* Critical time step min_e h_min(e)/c(e) (sld_updtss.f90:119-162). */
double sld_dev_critical_dt(const sld_dev *dev);

/* Field copies (synchronise the stream).  Element fields come back in the
 * caller's element order. */
#define SLD_FIELD_U        0   /* [npoin][3] displacement */
#define SLD_FIELD_V        1   /* [npoin][3] velocity */
#define SLD_FIELD_A        2   /* [npoin][3] acceleration */
#define SLD_FIELD_FINT     3   /* [npoin][3] assembled internal force */
#define SLD_FIELD_MASS     4   /* [npoin] lumped mass */
#define SLD_FIELD_SDV      5   /* [nelem][8][3] history ITER_K */
#define SLD_FIELD_STRESS   6   /* [nelem][8][9] PK2 (needs sld_dev_set_store) */
#define SLD_FIELD_FLAGS    7   /* [nelem][8] uint32 (needs sld_dev_set_store) */
#define SLD_FIELD_TIME     8   /* [1] current time */
#define SLD_FIELD_COUNTERS 9   /* [8] uint64 Gauss-point counts per flag bit 0..6 */
#define SLD_FIELD_PK1      10  /* [nelem][8][9] first Piola P = F S, column-major
                                  (gppio_sld, sld_builtin_materials.f90:355-371;
                                  needs sld_dev_set_store) */
#define SLD_FIELD_REACTIONS 11 /* [npoin][3] frxid_sld: the element RHS f_ext - f_int
                                  on prescribed dofs, 0 elsewhere
                                  (sld_elmmat.f90:228-237; no element body forces
                                  here, so -f_int); Alya's node-set FRXIX is its
                                  negative (sld_nodset.f90:100) */
#define SLD_FIELD_HEALTH   12  /* [4] run health: {halt bits (1 = det F <= 0 at a Gauss
                                  point, sld_elmpre.f90:229-240 -> sld_detjac.f90:93;
                                  2 = NaN or infinite displacement or residual,
                                  sld_cvgunk.f90:195-199), element of the most
                                  negative det F (caller numbering) or -1, that
                                  det F (FP32 precision) or 0, lowest node with a
                                  non-finite value or -1} */
int sld_dev_get(sld_dev *dev, int32_t field, void *host, int64_t count);

/* Overwrite state (restart, sld_restar.f90:63-65): fields U, V, A, SDV
 * (both history slots) and TIME.  Setting any of them clears a halt. */
int sld_dev_set(sld_dev *dev, int32_t field, const void *host, int64_t count);

/* Colouring used by the handle: colour_ptr [ncolor+1] and elem_order
 * [nelem], the caller's elements sorted by colour, colour c =
 * elem_order[colour_ptr[c] .. colour_ptr[c+1]-1]; within a colour in the
 * caller's order, except that a mesh mixing constitutive laws lists its SMA
 * elements before its neo-Hookean ones (the handle stores elements grouped
 * by law).  Both assemblies add the element forces at a node in this
 * order. */
int sld_dev_coloring(const sld_dev *dev, int32_t *ncolor, int32_t *color_ptr,
                     int32_t max_colors, int32_t *elem_order);

/* Launch shape of the element kernel: elements per workgroup (8, 16 or 32;
 * default 8 = one wave per workgroup), XCD-contiguous tile mapping on/off
 * (default on), persistent = workgroups per CU of a persistent launch
 * (0 = one workgroup per tile, default).  sld_dev_autotune picks these by
 * timing. */
int sld_dev_set_tile(sld_dev *dev, int32_t elems_per_block, int32_t xcd_remap, int32_t persistent);

/* Assembly of the internal force (both deterministic, bitwise equal):
 * SLD_ASSEMBLY_SCATTER = one element launch per colour adding into f_int
 * (the coloured element loop of sld_elmope_all, sld_element_operations.f90
 * with mod_matrix.f90:2327-2431 assembly); SLD_ASSEMBLY_GATHER (default)
 * = one element launch over all elements writing per-corner forces, summed
 * per node in colour order by the corrector.  Meshes with an element that
 * lists a node twice only accept GATHER. */
#define SLD_ASSEMBLY_SCATTER 0
#define SLD_ASSEMBLY_GATHER  1
int sld_dev_set_assembly(sld_dev *dev, int32_t assembly);

/* Neumann loads on boundary faces (replaces sld_bouope,
 * sld_bouope.f90:19-386, with the per-step boundary values of
 * sld_updbcs.f90:153-181).  Faces are QUA04 (4 nodes in the order of the
 * reference's shape2: (s,t) = (-1,-1), (1,-1), (1,1), (-1,1)); each is
 * integrated at 2x2 Gauss points (+-0.577350269189626, weight 1).  At
 * t^{n+1} a face carries
 *   code 3 (pressure):  traction = val[0] n, n = t_s x t_t / |t_s x t_t|
 *   code 6 (traction):  traction = val[0..2]
 * with val = value[f][0..2] under constant boundary conditions
 * (transient = 0), else (column x of function func[f] at t^{n+1}, 0, 0), or 0
 * where func is NULL or -1 (sld_funbou.f90 itask 1: the reference's
 * transient boundary values).  follower = 1 integrates over the current
 * faces X + u^{n+1} (FOLLOWER_LOADS: ON, kfl_follo_sld), else over the
 * reference faces.  Loads enter f_ext of dofs that are not prescribed;
 * the corrector adds each node's face corners in ascending face order.
 * nface = 0 removes the faces.  Functions are those of sld_bc_desc. */
typedef struct {
  int64_t nface;
  const int32_t *lnodb;        /* [nface][4] node indices (0-based) */
  const int32_t *code;         /* [nface] 3 or 6 (kfl_fixbo_sld) */
  const double *value;         /* [nface][3] bvnat_sld */
  const int32_t *func;         /* [nface] function index or -1 (kfl_funbo_sld), or NULL */
  int32_t transient;           /* 1: BOUNDARY_CONDITIONS, TRANSIENT (kfl_conbc_sld = 0) */
  int32_t follower;            /* 1: FOLLOWER_LOADS: ON */
} sld_face_desc;
int sld_dev_set_faces(sld_dev *dev, const sld_face_desc *faces);

/* Gravity (GRAVITY: NORM, GX, GY, GZ of the .sld.dat, sld_reaphy.f90:209-223;
 * its element load in sld_elmmat.f90:152-157): g[3] = grnor_sld *
 * gravi_sld, the acceleration (the caller normalises the direction).  Every
 * dof, prescribed ones included, receives m_a g, which is the element load
 * sum_g N_a rho g w detJ summed over the elements at the node (the lumped
 * mass row sum); the reactions field includes it.  g = NULL removes it. */
int sld_dev_set_gravity(sld_dev *dev, const double *g);

/* Body force (BODYF = func in the .sld.dat boundary conditions,
 * sld_reabcs.f90:523; its element load in sld_elmmat.f90:156-157 with
 * bodyf_sld from sld_updbcs.f90:96-99): every dof, prescribed ones included,
 * receives m_a b(t^{n+1}), where b is load function func evaluated by the
 * rule of sld_funbou.f90 itask 3 (linear between the rows with
 * t1 <= t < t2, zero outside the table) and m_a the lumped mass (the
 * element load summed over the node's elements).  Adds to gravity.
 * func = -1 removes it. */
int sld_dev_set_body_force(sld_dev *dev, int32_t func);

/* Rayleigh damping (RAYLEIGH_DAMPING: RAYALPHA, RAYBETA of the .sld.dat
 * material, dampi_sld(1:2, imate), sld_reaphy.f90:1046-1049), the element
 * terms of sld_elmmat.f90:160-163 and :174-194: the residual of node a
 * loses alpha sum_g N_a rho w detJ v(g) (v interpolated from
 * veloc_sld(ITER_K): v^n under central differences, the predicted v^{n+1}
 * under Tchamwa-Wielgosz) and gains, in each of its three components, the
 * beta term as written, beta sum_g sum_jk dN_a/dX_j L_kj S_jk w detJ (L the
 * velocity gradient, S the PK2 stress).  alpha_beta [nmat][2] (>= 0);
 * NULL or all zero removes it.  The internal force field stays f_int
 * alone (finte_sld); the reactions include the damping force.  Damping
 * runs with the gather assembly only (selected here). */
int sld_dev_set_damping(sld_dev *dev, const double *alpha_beta);

/* Energy balance of the explicit run (replaces sld_energy,
 * mod_sld_energy.f90:71-162, called from sld_endste.f90): on = 1 starts a
 * fresh balance (allocated on first use), on = 0 stops accumulating.  Each
 * step adds the trapezoidal work of the internal and external forces over
 * the displacement increment (full weight on the first step of a balance,
 * half afterwards, absolute value of each step's sum; the previous step's
 * forces are zero at the start, as fintt/fextt_sld(TIME_N)) and sets the kinetic
 * energy 0.5 sum m v.v from the corrected velocity.  Sums are fixed-order:
 * repeat runs are bitwise equal.  out[5] = {allie, allwk, allke, etota =
 * allie + allke - allwk, steps accumulated}. */
int sld_dev_set_energy(sld_dev *dev, int32_t on);
int sld_dev_energy(sld_dev *dev, double *out);

/* Autotuner: times nsteps eager steps for every launch shape (assembly
 * scatter/gather x elements per workgroup 8/16/32 x XCD mapping on/off x
 * persistent 0/2/4 workgroups per CU) from the same saved state, restores
 * the state, keeps the fastest.  best[4] = (elems_per_block, xcd_remap,
 * persistent, assembly); best_ms = ms/step. */
int sld_dev_autotune(sld_dev *dev, double dt, int32_t scheme, double phi, int32_t nsteps,
                     int32_t *best, double *best_ms);

/* Stream the handle enqueues on (a hipStream_t). */
void *sld_dev_stream(sld_dev *dev);

/* Wait for all work of the handle. */
int sld_dev_sync(sld_dev *dev);

int sld_dev_destroy(sld_dev *dev);

/* ------------------------------------------------------------------------
 * MLP surrogate of the return map (no counterpart in the reference; its
 * closest analogue is the per-Gauss-point ANN hook ann_framework%eval,
 * alya-ann/mod_ann_framework.f90:489-556).  16 -> 128 x4 -> 12, SiLU,
 * BF16 weights and activations, FP32 accumulation on MFMA.
 * Parameters in the canonical FP32 layout: W_l [n_out][n_in] for the five
 * layers back to back (52736 floats), then the biases (524 floats).
 * Feature and target definitions: alya_sma_amd/csrc/surrogate.hpp.
 * ------------------------------------------------------------------------ */
typedef struct sma_mlp sma_mlp;
#define SMA_MLP_NIN      16
#define SMA_MLP_NOUT     12
#define SMA_MLP_NWEIGHT  52736
#define SMA_MLP_NBIAS    524

/* Allocate and He-initialise (seeded) on the current device. */
int sma_mlp_create(uint64_t seed, sma_mlp **out);
int sma_mlp_destroy(sma_mlp *mlp);
/* Host FP32 canonical parameters in / out. */
int sma_mlp_set_params(sma_mlp *mlp, const float *W, const float *b);
int sma_mlp_get_params(sma_mlp *mlp, float *W, float *b);
/* Per-output weights w_j of the loss (1/n) sum w_j (y_j - t_j)^2 (12 floats). */
int sma_mlp_set_loss_weights(sma_mlp *mlp, const float *w12);
/* Fused inference: X [n][16] -> Y [n][12], device pointers, async. */
int sma_mlp_forward_device(sma_mlp *mlp, int64_t n, const float *X, float *Y, void *stream);
/* Loss and gradient (canonical layout, device pointers) of a batch. */
int sma_mlp_grad_device(sma_mlp *mlp, int64_t n, const float *X, const float *T,
                        float *gW, float *gb, float *loss_out, void *stream);
/* One Adam step on a batch (FP32 master weights, BF16 copies refreshed). */
int sma_mlp_train_step_device(sma_mlp *mlp, int64_t n, const float *X, const float *T,
                              double lr, float *loss_out, void *stream);
/* Training samples from the exact FP64 return map: X [n][16], T [n][12],
 * flags [n]; props_dev = base material [18] (device). */
int sma_mlp_gen_samples_device(int64_t n, uint64_t seed, const double *props_dev,
                               double strain_sd, double step_sd, double jitter,
                               float *X, float *T, uint32_t *flags, void *stream);

/* Attach a trained surrogate to a step handle (mode SLD_MODE_SURROGATE):
 * single-branch transforming Gauss points take the network's transformation
 * update x when it lies within tol (in xS) of the return map's solution --
 * |R(x)/R'(x)| of the reference's Newton residual, exact because that
 * residual is linear in the increment on one branch -- and the inputs are
 * in the trained range; otherwise the exact return map runs (counted as
 * SMA_FLAG_SURR_FALLBACK).  NULL detaches. */
int sld_dev_set_surrogate(sld_dev *dev, sma_mlp *mlp, double tol);

/* ------------------------------------------------------------------------
 * Device memory helpers for hosts without their own HIP allocator (the
 * Python and Fortran hosts use these for device-pointer entry points).
 * kind: 0 host->device, 1 device->host, 2 device->device.
 * ------------------------------------------------------------------------ */
int sma_device_malloc(void **p, int64_t bytes);
int sma_device_free(void *p);
int sma_memcpy(void *dst, const void *src, int64_t bytes, int32_t kind);
int sma_device_synchronize(void);
int sma_device_count(int32_t *n);
/* Streaming device-to-device copy kernel of `bytes` (multiple of 16): the
 * calibration point for HBM byte counters (FETCH_SIZE / WRITE_SIZE). */
int sma_copy_probe(const void *src, void *dst, int64_t bytes, void *stream);

/* ------------------------------------------------------------------------
 * Host helpers (no GPU needed).
 * ------------------------------------------------------------------------ */
/* Structured box of nx*ny*nz HEX08 over [0,lx]x[0,ly]x[0,lz]; node (i,j,k)
 * has index i + (nx+1)*(j + (ny+1)*k); element (i,j,k) index
 * i + nx*(j + ny*k).  coord [(nx+1)(ny+1)(nz+1)][3], lnods [nx*ny*nz][8]. */
int sld_mesh_box(int32_t nx, int32_t ny, int32_t nz, double lx, double ly, double lz,
                 double *coord, int32_t *lnods);

/* Greedy element colouring (no two elements of one colour share a node),
 * elements visited in index order.  color [nelem] out; returns ncolor or < 0. */
int32_t sld_color_elements(int64_t nelem, int64_t npoin, const int32_t *lnods,
                           int32_t *color);

#ifdef __cplusplus
}
#endif
#endif /* ALYA_SMA_H */
