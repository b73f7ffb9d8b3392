// C ABI of the MI355X explicit SMA step: handle management, colouring, the
// per-step launch sequence (captured into hipGraphs) and the point-wise law.
// Declarations and the reference procedures each entry point replaces are in
// include/alya_sma.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "alya_sma.h"
#include "common.hpp"
#include "sld_kernels.hpp"

namespace {

using sma_abi::fail;
#define HIPCHK(x) SMA_HIPCHK(x)

template <typename T>
int dalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  HIPCHK(hipMalloc((void**)p, count * sizeof(T)));
  return 0;
}

// Synchronous upload (set-up only): the host buffer may go away on return.
template <typename T>
int upload(T** p, const T* host, size_t count, hipStream_t) {
  int rc = dalloc(p, count);
  if (rc) return rc;
  if (count) HIPCHK(hipMemcpy(*p, host, count * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

unsigned nblocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// Workgroup size of the corrector without the energy balance.
#ifndef SLD_CORRECT_THREADS
#define SLD_CORRECT_THREADS 64
#endif

}  // namespace

// Steps per captured graph of a long run, longest first (even lengths: see
// sld_dev_step); each run pays one stand-alone predictor.
constexpr int GRAPH_RUNS[] = {64, 16, 2};

struct GraphKey {
  double dt, phi;
  int scheme, mode, store, epb, xcd, persistent, assembly, tn, energy, len, damp, body;
  bool operator<(const GraphKey& o) const {
    return std::tie(dt, phi, scheme, mode, store, epb, xcd, persistent, assembly, tn, energy, len,
                    damp, body) <
           std::tie(o.dt, o.phi, o.scheme, o.mode, o.store, o.epb, o.xcd, o.persistent,
                    o.assembly, o.tn, o.energy, o.len, o.damp, o.body);
  }
};

struct sld_dev {
  int device = 0;
  hipStream_t stream = nullptr;
  int64_t npoin = 0, nelem = 0;
  int32_t ncolor = 0;
  std::vector<int32_t> color_ptr;   // [ncolor+1] ranges of color_list
  std::vector<int32_t> color_list;  // stored elements sorted by colour (stable)
  std::vector<int32_t> order;       // stored position -> caller element
  int32_t* color_list_dev = nullptr;
  // device arrays
  int32_t* lnods = nullptr;
  double* coord[3] = {};
  double* u[3] = {};
  double* v[3] = {};
  double* a[3] = {};
  double* fint64[3] = {};
  float* fint32[3] = {};
  // gather assembly: corner forces in node-sorted slots (FP64 size, reused
  // for FP32), slot ranges per node, slot of every element corner
  void* efor = nullptr;
  int32_t* node_ptr = nullptr;
  int32_t* corner_slot = nullptr;
  bool repeated_nodes = false;       // an element lists a node twice: gather only
  int32_t slot_width = 0;            // corner slots per node when padded (gather), else 0
  size_t efor_bytes = 0;
  double* mass = nullptr;
  double* rmass = nullptr;
  uint8_t* fixmask = nullptr;
  int32_t* func_of_node = nullptr;
  double* func_table = nullptr;
  int32_t* func_npts = nullptr;
  int32_t* func_smooth = nullptr;
  int32_t nfunc = 0, func_maxpts = 0;
  double* fext0[3] = {};
  int32_t fext_func = -1;
  double* props = nullptr;
  double* thresholds = nullptr;      // [nmat][4] Rf_AS, Rs_AS, Rf_SA, Rs_SA
  int32_t nmat = 0;
  double* rho = nullptr;
  int32_t* mat_of_elem = nullptr;
  double* E_of_elem = nullptr;
  double* kg = nullptr;              // [nelem][2] bulk, shear modulus
  uint8_t* law_of_elem = nullptr;    // [nelem] 1 = neo-Hookean; null when all SMA
  int32_t nh_begin = -1;             // first stored neo-Hookean element (law-grouped), or -1
  double* sdv[2][3] = {};
  int cur = 1;                       // index of the TIME_N history and clock slot
  double* S_out = nullptr;
  double* P_out = nullptr;
  uint32_t* flags_out = nullptr;
  unsigned long long* counters = nullptr;
  double* clock = nullptr;           // [2] t^n in slot cur
  sld::Health* health = nullptr;     // run health: set by the kernel that finds a failure
  unsigned halted = 0;               // host copy of health->halted (last synchronising call)
  std::string halt_msg;
  // energy balance, allocated when first switched on
  int energy = 0;
  double* du[3] = {};
  double* fint_prev[3] = {};
  double* fext_prev[3] = {};
  double* energy_part = nullptr;     // [3][corrector blocks]
  double* energy_acc = nullptr;      // allie, allwk, allke, etota, steps
  // boundary faces with Neumann loads (sld_dev_set_faces)
  struct Faces {
    int64_t nface = 0;
    int32_t transient = 0, follower = 0;
    int32_t* lnodb = nullptr;
    int32_t* code = nullptr;
    double* value = nullptr;
    int32_t* func = nullptr;
    int32_t* slot = nullptr;           // [4 nface] node-sorted slot of each corner
    int32_t* node_ptr = nullptr;       // [npoin+1]
    double* fface = nullptr;           // [4 nface][3]
  } faces;
  int mode = SLD_MODE_FP64, store = 0, epb = 8, xcd = 1, persistent = 0, use_graph = 1;
  int gravity = 0;                   // sld_dev_set_gravity
  double grav[3] = {0.0, 0.0, 0.0};
  int32_t body_func = -1;            // sld_dev_set_body_force
  std::vector<double> func_table_host;   // host copies of the load tables
  std::vector<int32_t> func_npts_host;
  // Rayleigh damping (sld_dev_set_damping), allocated on first use
  int damp = 0;
  double* damping = nullptr;         // [nmat][2] (alpha, beta)
  double* vdamp[3] = {};             // v^n kept by the central-difference predictor
  void* edamp = nullptr;             // corner damping forces, laid out as efor
  double* fdamp[3] = {};             // nodal damping force of the last step
  int assembly = SLD_ASSEMBLY_GATHER;
  double dtcrit = 0.0;
  std::map<GraphKey, std::pair<hipGraph_t, hipGraphExec_t>> graphs;
  std::vector<void*> allocs;
  struct Sur {
    sma_mlp* mlp = nullptr;
    double tol = 1e-4;
    double T = 310.0;
    int64_t seg_cap = 0;               // slots per segment of the list
    int64_t cap = 0;                   // slots of the list (LIST_SEGS x seg_cap)
    uint8_t* status = nullptr;         // [gp] 0 not flagged, 1 listed, 2 flagged, not listed
    int32_t* slot = nullptr;           // [gp] slot of a listed point
    float* feat = nullptr;             // [cap][16]
    unsigned* count = nullptr;         // segment counters, LIST_STRIDE apart
    float* Y = nullptr;
    // candidate elements by history slot: cand[tn] is read by the prepass
    // of the step from slot tn, cand[1 - tn] written by its element launch
    uint8_t* cand[2] = {nullptr, nullptr};
  };
  Sur* sur = nullptr;                // attached surrogate (mode 2)
};

namespace {

template <typename T>
int track(sld_dev* d, T* p) {
  d->allocs.push_back((void*)p);
  return 0;
}

sld::NodalArgs nodal_args(sld_dev* d, int tn, int scheme = 0) {
  sld::NodalArgs A{};
  A.npoin = d->npoin;
  const bool f32 = d->mode != SLD_MODE_FP64;
  const bool gather = d->assembly == SLD_ASSEMBLY_GATHER;
  for (int k = 0; k < 3; ++k) {
    A.u[k] = d->u[k];
    A.v[k] = d->v[k];
    A.a[k] = d->a[k];
    A.fint[k] = f32 ? (void*)d->fint32[k] : (void*)d->fint64[k];
    A.fext0[k] = d->fext0[k];
  }
  A.efor = gather ? d->efor : nullptr;
  A.tn = tn;
  A.zero_fint = gather ? 0 : 1;
  A.node_ptr = gather ? d->node_ptr : nullptr;
  A.slot_width = gather ? d->slot_width : 0;
  A.sur_count = d->sur ? d->sur->count : nullptr;
  A.fint_fp32 = f32 ? 1 : 0;
  A.rmass = d->rmass;
  A.fixmask = d->fixmask;
  A.func_of_node = d->func_of_node;
  A.func_table = d->func_table;
  A.func_npts = d->func_npts;
  A.func_smooth = d->func_smooth;
  A.func_maxpts = d->func_maxpts;
  A.fext_func = d->fext_func;
  A.clock = d->clock;
  if (d->energy) {
    for (int k = 0; k < 3; ++k) {
      A.du[k] = d->du[k];
      A.fint_prev[k] = d->fint_prev[k];
    }
    for (int k = 0; k < 3; ++k) A.fext_prev[k] = d->fext_prev[k];
    A.mass = d->mass;
    A.energy_part = d->energy_part;
  }
  if (d->faces.nface > 0) {
    A.node_face_ptr = d->faces.node_ptr;
    A.fface = d->faces.fface;
  }
  if (d->gravity) {
    A.gravity = 1;
    A.mass = d->mass;
    for (int k = 0; k < 3; ++k) A.grav[k] = d->grav[k];
  }
  A.body_func = d->body_func;
  if (d->body_func >= 0) A.mass = d->mass;
  if (d->damp) {
    for (int k = 0; k < 3; ++k) {
      A.vdamp[k] = scheme == 0 ? d->vdamp[k] : nullptr;
      A.fdamp[k] = d->fdamp[k];
    }
    A.edamp = d->edamp;
  }
  A.health = d->health;
  return A;
}

// Clears the run health (creation and every restart from a set state).
int clear_health(sld_dev* d) {
  const sld::Health h{0ull, ~0ull, 0u, 0u};
  HIPCHK(hipMemcpy(d->health, &h, sizeof(h), hipMemcpyHostToDevice));
  d->halted = 0;
  d->halt_msg.clear();
  return 0;
}

// Reads the run health after the handle's stream has drained.  Returns
// SMA_ERR_STATE (message naming the element or node) when the run halted.
int read_health(sld_dev* d, sld::Health* out = nullptr) {
  sld::Health h{};
  HIPCHK(hipMemcpy(&h, d->health, sizeof(h), hipMemcpyDeviceToHost));
  if (out) *out = h;
  d->halted = h.halted;
  if (!h.halted) return 0;
  std::string msg = "run halted:";
  if (h.halted & sld::HALT_DETF) {
    const uint32_t lo = (uint32_t)(h.detf & 0xffffffffull);
    const int64_t e = (int64_t)(0xffffffffu - lo);
    uint32_t bits = (uint32_t)(h.detf >> 32);
    float mag;
    std::memcpy(&mag, &bits, sizeof(mag));
    const int64_t ec = (e >= 0 && e < d->nelem) ? d->order[e] : e;
    msg += " negative Jacobian det F = " + std::to_string(-(double)mag) + " at element " +
           std::to_string(ec) + " (sld_detjac)";
  }
  if (h.halted & sld::HALT_NONFINITE)
    msg += " NaN or infinite displacement or residual at node " + std::to_string(h.node) +
           " (sld_cvgunk)";
  d->halt_msg = msg;
  return fail(SMA_ERR_STATE, msg);
}

int g_num_cu = 0;

template <typename R, int EPB, bool STORE, int SUR, bool GATHER, int LAWS = 0, bool DAMP = false>
void launch_element(sld::ElemArgs A, hipStream_t s, int persistent) {
  const int ne = A.e_end - A.e_begin;
  if (ne <= 0) return;
  const int ntiles = (int)nblocks(ne, EPB);
  int grid = ntiles;
  A.tiles_per_block = 0;
  if (persistent > 0) {
    // persistent: `persistent` workgroups per CU, each walking a run of tiles
    const int cap = std::max(1, g_num_cu) * persistent;
    if (ntiles > cap) {
      A.tiles_per_block = (ntiles + cap - 1) / cap;
      grid = (ntiles + A.tiles_per_block - 1) / A.tiles_per_block;
    }
  }
  hipLaunchKernelGGL((sld::sld_element<R, EPB, STORE, SUR, GATHER, LAWS, DAMP>), dim3(grid),
                     dim3(EPB * 8), 0, s, A);
}

// Damping kernels come in one tile shape (8 elements per workgroup).
template <typename R, bool STORE, int SUR, bool GATHER, int LAWS = 0, bool DAMP = false>
void launch_element_epb(int epb, const sld::ElemArgs& A, hipStream_t s, int persistent) {
  if constexpr (DAMP) {
    launch_element<R, 8, STORE, SUR, GATHER, LAWS, true>(A, s, persistent);
  } else {
    if (epb == 8) launch_element<R, 8, STORE, SUR, GATHER, LAWS>(A, s, persistent);
    else if (epb == 16) launch_element<R, 16, STORE, SUR, GATHER, LAWS>(A, s, persistent);
    else launch_element<R, 32, STORE, SUR, GATHER, LAWS == 2 ? 0 : LAWS>(A, s, persistent);
  }
}

// A launch whose elements all take the SMA law (no law_of_elem, or the SMA
// range of a law-grouped mesh) runs the kernel without the neo-Hookean path.
template <typename R, bool STORE, int SUR, bool GATHER, bool DAMP = false>
void launch_element_laws(const sld_dev* d, const sld::ElemArgs& A, hipStream_t s) {
  if (d->law_of_elem) launch_element_epb<R, STORE, SUR, GATHER, 0, DAMP>(d->epb, A, s, d->persistent);
  else launch_element_epb<R, STORE, SUR, GATHER, 1, DAMP>(d->epb, A, s, d->persistent);
}

template <typename R, int SUR, bool DAMP>
void launch_element_damp(const sld_dev* d, const sld::ElemArgs& A, hipStream_t s) {
  const bool st = d->store != 0;
  if constexpr (SUR != 1) {
    if (d->nh_begin >= 0 && A.e_end > d->nh_begin) {
      // law-grouped mesh: the SMA range and the neo-Hookean range each with
      // the kernel that has only its law (surrogate mode: the SMA range in
      // its surrogate kernel, the neo-Hookean range, which no network
      // serves, in the plain one)
      sld::ElemArgs B = A;
      B.e_end = std::min(A.e_end, d->nh_begin);
      if (B.e_end > B.e_begin) {
        if (st) launch_element_epb<R, true, SUR, true, 1, DAMP>(d->epb, B, s, d->persistent);
        else launch_element_epb<R, false, SUR, true, 1, DAMP>(d->epb, B, s, d->persistent);
      }
      B.e_begin = std::max(A.e_begin, d->nh_begin);
      B.e_end = A.e_end;
      if (st) launch_element_epb<R, true, 0, true, 2, DAMP>(d->epb, B, s, d->persistent);
      else launch_element_epb<R, false, 0, true, 2, DAMP>(d->epb, B, s, d->persistent);
      return;
    }
  }
  if (st) launch_element_laws<R, true, SUR, true, DAMP>(d, A, s);
  else launch_element_laws<R, false, SUR, true, DAMP>(d, A, s);
}

template <typename R, int SUR>
void launch_element_mode(const sld_dev* d, const sld::ElemArgs& A, hipStream_t s, bool gather) {
  const bool st = d->store != 0;
  if (gather) {
    // damping runs with the gather assembly only (its corner forces)
    if (d->damp) launch_element_damp<R, SUR, true>(d, A, s);
    else launch_element_damp<R, SUR, false>(d, A, s);
    return;
  }
  if (st) launch_element_laws<R, true, SUR, false>(d, A, s);
  else launch_element_laws<R, false, SUR, false>(d, A, s);
}

bool surrogate_attached(const sld_dev* d) { return d->sur != nullptr && d->sur->mlp != nullptr; }

// Surrogate state for a step that does not follow the previous one: after a
// state set from outside the stepping (restart, autotune trials, clearing a
// halt), a mode switch or attaching a network, every element is a candidate
// of the next prepass (the previous step's transforming elements are not
// known) and the list starts empty (a step that halted before its
// corrector leaves its counts behind).
int surrogate_restart(sld_dev* d) {
  if (!d->sur) return SMA_OK;
  for (int k = 0; k < 2; ++k)
    if (d->sur->cand[k]) HIPCHK(hipMemsetAsync(d->sur->cand[k], 1, (size_t)d->nelem, d->stream));
  if (d->sur->count)
    HIPCHK(hipMemsetAsync(d->sur->count, 0, mlp::LIST_SEGS * mlp::LIST_STRIDE * sizeof(unsigned),
                          d->stream));
  HIPCHK(hipStreamSynchronize(d->stream));
  return SMA_OK;
}

void surrogate_detach(sld_dev* d) {
  if (!d->sur) return;
  void* ps[] = {d->sur->status, d->sur->slot, d->sur->feat, d->sur->count, d->sur->Y,
                d->sur->cand[0], d->sur->cand[1]};
  for (void* p : ps) if (p) (void)hipFree(p);
  delete d->sur;
  d->sur = nullptr;
}

// Optional timing (eager launches only, never captured): event pairs
// recorded around the predictor, the element phase (all its launches), the
// corrector and the surrogate's extra launches.  With the gather assembly
// the element launch is issued elem_reps times back to back (it rewrites
// the same outputs from the same inputs; the repeats count flags into
// scratch_counters), so the bracket measures launches, not event overhead.
struct KernelTimer {
  std::vector<hipEvent_t> ev;
  std::vector<int> kind;        // 0 predict, 1 element phase, 2 correct, 3-4 surrogate
                                // prepass and inference
  size_t used = 0;
  int element_launches = 0;
  int elem_reps = 1;
  unsigned long long* scratch_counters = nullptr;
  hipStream_t s = nullptr;
  void mark(int k) {
    if (used + 2 > ev.size()) {
      hipEvent_t a, b;
      hipEventCreate(&a);
      hipEventCreate(&b);
      ev.push_back(a);
      ev.push_back(b);
      kind.push_back(k);
    } else {
      kind[used / 2] = k;
    }
    hipEventRecord(ev[used], s);
  }
  void done() {
    hipEventRecord(ev[used + 1], s);
    used += 2;
  }
  ~KernelTimer() {
    for (auto e : ev) hipEventDestroy(e);
  }
};

#define TBEGIN(k) do { if (T) T->mark(k); } while (0)
#define TEND() do { if (T) T->done(); } while (0)

// One explicit step; tn = history slot holding TIME_N.  predict = false:
// the previous step's corrector already ran this step's predictor (fused);
// fuse = true: this step's corrector runs the next step's predictor.
int enqueue_step(sld_dev* d, double dt, int scheme, double phi, int tn, KernelTimer* T = nullptr,
                 bool predict = true, bool fuse = false) {
  hipStream_t s = d->stream;
  sld::NodalArgs N = nodal_args(d, tn, scheme);
  const bool gather = d->assembly == SLD_ASSEMBLY_GATHER;
  const unsigned nb = nblocks(d->npoin, 256);
  TBEGIN(0);
  if (predict) {
    if (scheme == 0) hipLaunchKernelGGL((sld::sld_predict<0>), dim3(nb), dim3(256), 0, s, N, dt, phi);
    else hipLaunchKernelGGL((sld::sld_predict<1>), dim3(nb), dim3(256), 0, s, N, dt, phi);
  }
  if (d->faces.nface > 0) {
    const sld_dev::Faces& Fc = d->faces;
    sld::FaceArgs F{};
    F.nface = Fc.nface;
    F.lnodb = Fc.lnodb;
    F.code = Fc.code;
    F.value = Fc.value;
    F.func = Fc.func;
    F.transient = Fc.transient;
    F.follower = Fc.follower;
    for (int k = 0; k < 3; ++k) {
      F.coord[k] = d->coord[k];
      F.u[k] = d->u[k];
    }
    F.fixmask = d->fixmask;
    F.func_table = d->func_table;
    F.func_npts = d->func_npts;
    F.func_smooth = d->func_smooth;
    F.func_maxpts = d->func_maxpts;
    F.slot = Fc.slot;
    F.fface = Fc.fface;
    F.clock = d->clock;
    F.tn = tn;
    hipLaunchKernelGGL(sld::sld_face_load, dim3(nblocks(Fc.nface, 256)), dim3(256), 0, s, F, dt);
  }
  TEND();

  {
    sld::ElemArgs E{};
    E.nelem = d->nelem;
    E.lnods = d->lnods;
    for (int k = 0; k < 3; ++k) {
      E.coord[k] = d->coord[k];
      E.u[k] = d->u[k];
      E.sdv_n[k] = d->sdv[tn][k];
      E.sdv_k[k] = d->sdv[1 - tn][k];
      E.fint[k] = N.fint[k];
    }
    E.efor = d->efor;
    E.corner_slot = d->corner_slot;
    E.props = d->props;
    E.thresholds = d->thresholds;
    E.mat_of_elem = d->mat_of_elem;
    E.E_of_elem = d->E_of_elem;
    E.kg = d->kg;
    E.law_of_elem = d->law_of_elem;
    E.S_out = d->S_out;
    E.P_out = d->P_out;
    E.flags_out = d->flags_out;
    E.counters = d->counters;
    E.health = d->health;
    if (d->damp) {
      // veloc_sld(ITER_K) at the element loop: v^n (central difference,
      // kept by the predictor) or the predicted v^{n+1} (Tchamwa-Wielgosz)
      for (int k = 0; k < 3; ++k) E.vel[k] = scheme == 0 ? d->vdamp[k] : d->v[k];
      E.damping = d->damping;
      E.rho = d->rho;
      E.edamp = d->edamp;
    }
    E.xcd_remap = d->xcd;
    const bool sur = d->mode == SLD_MODE_SURROGATE;
    if (sur) {
      // prepass over the candidate elements (those with a transforming
      // point in the previous step): flags and features of their
      // transforming points, listed for the network; the element launch
      // checks the network's answers (the other points take the exact map)
      sld_dev::Sur* S = d->sur;
      E.sur_cand = S->cand[tn];
      E.sur_status = S->status;
      E.sur_slot = S->slot;
      E.sur_feat = S->feat;
      E.sur_Y = S->Y;
      E.sur_tol = S->tol;
      E.sur_count = S->count;
      E.sur_seg_cap = S->seg_cap;
      E.T_feat = S->T;
      // a law-grouped mesh: the SMA range only (its kernel without the
      // neo-Hookean path)
      E.e_begin = 0;
      E.e_end = d->nh_begin >= 0 ? d->nh_begin : (int32_t)d->nelem;
      E.elem_list = nullptr;
      TBEGIN(3);
      if (d->nh_begin >= 0) {
        if (E.e_end > 0) launch_element_epb<float, false, 1, false, 1>(d->epb, E, s, d->persistent);
      } else {
        launch_element_laws<float, false, 1, false>(d, E, s);
      }
      TEND();
      TBEGIN(4);
      int rc = sma_mlp_infer_listed(S->mlp, S->feat, S->count, S->seg_cap, S->Y, s);
      TEND();
      if (rc) return rc;
      E.sur_cand = nullptr;
      E.sur_cand_next = S->cand[1 - tn];
    }
    // gather: one launch over every element; scatter: one launch per colour
    const int nlaunch = gather ? (T ? T->elem_reps : 1) : d->ncolor;
    TBEGIN(1);
    for (int c = 0; c < nlaunch; ++c) {
      E.e_begin = gather ? 0 : d->color_ptr[c];
      E.e_end = gather ? (int32_t)d->nelem : d->color_ptr[c + 1];
      E.elem_list = gather ? nullptr : d->color_list_dev;
      E.counters = (gather && c > 0) ? T->scratch_counters : d->counters;
      if (E.e_end <= E.e_begin) continue;
      if (T) T->element_launches++;
      if (d->mode == SLD_MODE_FP64) launch_element_mode<double, 0>(d, E, s, gather);
      else if (sur) launch_element_mode<float, 2>(d, E, s, gather);
      else launch_element_mode<float, 0>(d, E, s, gather);
    }
    TEND();
  }
  TBEGIN(2);
  {
    // small workgroups spread the nodes evenly over the CUs; the energy
    // balance keeps 256 (its partial sums are per 256-node block)
    const int cbs = d->energy ? 256 : SLD_CORRECT_THREADS;
    const dim3 g(nblocks(d->npoin, cbs)), b(cbs);
    if (fuse) {
      if (scheme == 0) hipLaunchKernelGGL((sld::sld_correct<0, true>), g, b, 0, s, N, dt, phi);
      else hipLaunchKernelGGL((sld::sld_correct<1, true>), g, b, 0, s, N, dt, phi);
    } else {
      if (scheme == 0) hipLaunchKernelGGL((sld::sld_correct<0, false>), g, b, 0, s, N, dt, phi);
      else hipLaunchKernelGGL((sld::sld_correct<1, false>), g, b, 0, s, N, dt, phi);
    }
  }
  if (d->energy)
    hipLaunchKernelGGL(sld::sld_energy_reduce, dim3(1), dim3(256), 0, s, d->energy_part, (int)nb,
                       d->energy_acc, (const sld::Health*)d->health);
  TEND();
  HIPCHK(hipGetLastError());
  return 0;
}

void free_graphs(sld_dev* d) {
  for (auto& kv : d->graphs) {
    hipGraphExecDestroy(kv.second.second);
    hipGraphDestroy(kv.second.first);
  }
  d->graphs.clear();
}

// Body-force table value (sld_funbou.f90:142-167, itask 3), as the corrector
// evaluates it on the device.
double host_funbou_body(const double* tab, int npts, int c, double t) {
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

double host_critical_dt(const double* coord, const int32_t* lnods, int64_t nelem,
                        const double* props, const double* rho, const int32_t* mat,
                        const double* Eel, const int32_t* law) {
  // sld_updtss.f90:119-162 with elmlen.f90 (ndime==3, hnatu = 2 for HEX08)
  static const int sg[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                               {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
  double dtmin = 1e300;
  for (int64_t e = 0; e < nelem; ++e) {
    double xm[3][3] = {{0}};
    for (int k = 0; k < 8; ++k) {
      const double* x = coord + 3 * (int64_t)lnods[8 * e + k];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) xm[a][b] += x[a] * 0.125 * sg[k][b];
    }
    double t1 = xm[1][1] * xm[2][2] - xm[2][1] * xm[1][2];
    double t2 = -xm[1][0] * xm[2][2] + xm[2][0] * xm[1][2];
    double t3 = xm[1][0] * xm[2][1] - xm[2][0] * xm[1][1];
    double det = xm[0][0] * t1 + xm[0][1] * t2 + xm[0][2] * t3;
    double dn = 1.0 / det, tr[3][3];
    tr[0][0] = t1 * dn; tr[1][0] = t2 * dn; tr[2][0] = t3 * dn;
    tr[1][1] = (xm[0][0] * xm[2][2] - xm[2][0] * xm[0][2]) * dn;
    tr[2][1] = (-xm[0][0] * xm[2][1] + xm[0][1] * xm[2][0]) * dn;
    tr[2][2] = (xm[0][0] * xm[1][1] - xm[1][0] * xm[0][1]) * dn;
    tr[0][1] = (-xm[0][1] * xm[2][2] + xm[2][1] * xm[0][2]) * dn;
    tr[0][2] = (xm[0][1] * xm[1][2] - xm[1][1] * xm[0][2]) * dn;
    tr[1][2] = (-xm[0][0] * xm[1][2] + xm[1][0] * xm[0][2]) * dn;
    double hmin = 1e300;
    for (int i = 0; i < 3; ++i)
      hmin = std::min(hmin, 2.0 / std::sqrt(tr[i][0] * tr[i][0] + tr[i][1] * tr[i][1] +
                                            tr[i][2] * tr[i][2]));
    const int m = mat ? mat[e] : 0;
    double c;
    if (law && law[m] == SLD_LAW_NEOHOOKEAN102) {
      // sld_velsnd.f90:75-78: sqrt((K0 + 4 mu0 / 3) / rho)
      c = std::sqrt((props[18 * m + 1] + 4.0 * props[18 * m] / 3.0) / rho[m]);
    } else {
      const double E = Eel ? Eel[e] : props[18 * m];
      const double nu = props[18 * m + 1];
      const double lam = nu * E / ((1.0 + nu) * (1.0 - 2.0 * nu));
      const double mu = E / (2.0 * (1.0 + nu));
      c = std::sqrt((lam + 2.0 * mu) / rho[m]);
    }
    dtmin = std::min(dtmin, hmin / c);
  }
  return dtmin;
}

}  // namespace


extern "C" {

const char* sma_last_error(void) { return sma_abi::last_error.c_str(); }

int32_t sld_color_elements(int64_t nelem, int64_t npoin, const int32_t* lnods, int32_t* color) {
  if (nelem < 0 || npoin <= 0 || (!lnods && nelem) || (!color && nelem))
    return fail(SMA_ERR_ARG, "sld_color_elements: bad arguments");
  std::vector<uint64_t> used((size_t)npoin, 0);
  int32_t ncolor = 0;
  for (int64_t e = 0; e < nelem; ++e) {
    uint64_t m = 0;
    for (int a = 0; a < 8; ++a) {
      const int32_t p = lnods[8 * e + a];
      if (p < 0 || p >= npoin) return fail(SMA_ERR_ARG, "sld_color_elements: node index out of range");
      m |= used[p];
    }
    if (m == ~0ull) return fail(SMA_ERR_ARG, "sld_color_elements: more than 64 colours needed");
    const int c = __builtin_ctzll(~m);
    color[e] = c;
    ncolor = std::max(ncolor, c + 1);
    for (int a = 0; a < 8; ++a) used[lnods[8 * e + a]] |= (1ull << c);
  }
  return ncolor;
}

int sld_mesh_box(int32_t nx, int32_t ny, int32_t nz, double lx, double ly, double lz,
                 double* coord, int32_t* lnods) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || !coord || !lnods)
    return fail(SMA_ERR_ARG, "sld_mesh_box: bad arguments");
  const int64_t px = nx + 1, py = ny + 1, pz = nz + 1;
  if (px * py * pz > INT32_MAX) return fail(SMA_ERR_ARG, "sld_mesh_box: too many nodes");
  for (int64_t k = 0; k < pz; ++k)
    for (int64_t j = 0; j < py; ++j)
      for (int64_t i = 0; i < px; ++i) {
        const int64_t p = i + px * (j + py * k);
        coord[3 * p + 0] = lx * (double)i / nx;
        coord[3 * p + 1] = ly * (double)j / ny;
        coord[3 * p + 2] = lz * (double)k / nz;
      }
  for (int64_t k = 0; k < nz; ++k)
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i) {
        const int64_t e = i + (int64_t)nx * (j + (int64_t)ny * k);
        auto P = [&](int64_t a, int64_t b, int64_t c) { return (int32_t)(a + px * (b + py * c)); };
        int32_t* L = lnods + 8 * e;
        L[0] = P(i, j, k);         L[1] = P(i + 1, j, k);
        L[2] = P(i + 1, j + 1, k); L[3] = P(i, j + 1, k);
        L[4] = P(i, j, k + 1);     L[5] = P(i + 1, j, k + 1);
        L[6] = P(i + 1, j + 1, k + 1); L[7] = P(i, j + 1, k + 1);
      }
  return SMA_OK;
}

int sld_dev_create(const sld_mesh_desc* mesh, const sld_material_desc* mat,
                   const sld_bc_desc* bc, sld_dev** out) {
  if (!mesh || !mat || !bc || !out) return fail(SMA_ERR_ARG, "sld_dev_create: null argument");
  *out = nullptr;
  const int64_t np = mesh->npoin, ne = mesh->nelem;
  if (np <= 0 || ne <= 0 || !mesh->lnods || !mesh->coord)
    return fail(SMA_ERR_ARG, "sld_dev_create: empty mesh");
  if (ne > (int64_t)INT32_MAX / 8 || np > INT32_MAX)
    return fail(SMA_ERR_ARG, "sld_dev_create: mesh too large for 32-bit indices");
  if (mat->nmat <= 0 || !mat->props || !mat->rho)
    return fail(SMA_ERR_ARG, "sld_dev_create: no material");
  if (!bc->fixmask) return fail(SMA_ERR_ARG, "sld_dev_create: fixmask required");
  if (mat->mat_of_elem)
    for (int64_t e = 0; e < ne; ++e)
      if (mat->mat_of_elem[e] < 0 || mat->mat_of_elem[e] >= mat->nmat)
        return fail(SMA_ERR_ARG, "sld_dev_create: material index out of range");
  bool any_neo = false;
  if (mat->law)
    for (int k = 0; k < mat->nmat; ++k) {
      if (mat->law[k] != SLD_LAW_SMA160 && mat->law[k] != SLD_LAW_NEOHOOKEAN102)
        return fail(SMA_ERR_ARG, "sld_dev_create: law must be SLD_LAW_SMA160 or SLD_LAW_NEOHOOKEAN102");
      any_neo |= mat->law[k] == SLD_LAW_NEOHOOKEAN102;
    }
  if (bc->nfunc < 0 || (bc->nfunc > 0 && (!bc->func_table || !bc->func_npts || bc->func_maxpts < 1)))
    return fail(SMA_ERR_ARG, "sld_dev_create: bad load-function tables");
  for (int f = 0; f < bc->nfunc; ++f)
    if (bc->func_npts[f] < 1 || bc->func_npts[f] > bc->func_maxpts)
      return fail(SMA_ERR_ARG, "sld_dev_create: func_npts out of range");
  if (bc->func_of_node)
    for (int64_t p = 0; p < np; ++p)
      if (bc->func_of_node[p] < -1 || bc->func_of_node[p] >= bc->nfunc)
        return fail(SMA_ERR_ARG, "sld_dev_create: func_of_node out of range");
  const int32_t fext_func = bc->fext0 ? bc->fext_func : -1;
  if (fext_func < -1 || fext_func >= bc->nfunc)
    return fail(SMA_ERR_ARG, "sld_dev_create: fext_func out of range");

  // colouring and colour-sorted order
  std::vector<int32_t> color((size_t)ne);
  const int32_t ncolor = sld_color_elements(ne, np, mesh->lnods, color.data());
  if (ncolor < 0) return ncolor;
  std::vector<int32_t> nadj((size_t)np + 1, 0);
  for (int64_t i = 0; i < 8 * ne; ++i) nadj[mesh->lnods[i] + 1]++;
  for (int64_t p = 0; p < np; ++p)
    if (!nadj[p + 1]) return fail(SMA_ERR_ARG, "sld_dev_create: node not used by any element (zero mass)");
  bool repeated = false;
  for (int64_t e = 0; e < ne && !repeated; ++e)
    for (int a = 0; a < 8; ++a)
      for (int b = a + 1; b < 8; ++b) repeated |= mesh->lnods[8 * e + a] == mesh->lnods[8 * e + b];

  sld_dev* d = new sld_dev();
  d->npoin = np;
  d->nelem = ne;
  d->ncolor = ncolor;
  d->repeated_nodes = repeated;
  d->color_ptr.assign(ncolor + 1, 0);
  for (int64_t e = 0; e < ne; ++e) d->color_ptr[color[e] + 1]++;
  for (int c = 0; c < ncolor; ++c) d->color_ptr[c + 1] += d->color_ptr[c];
  // elements are stored in the caller's order (neighbours stay close for
  // the gather launch), grouped by constitutive law when the mesh mixes
  // laws (SMA first, the caller's order within each law) so that the eight
  // elements of a wave take one branch; the coloured scatter walks them
  // through color_list (stored positions)
  d->order.resize(ne);
  for (int64_t e = 0; e < ne; ++e) d->order[e] = (int32_t)e;
  if (any_neo)
    std::stable_partition(d->order.begin(), d->order.end(), [&](int32_t e) {
      return mat->law[mat->mat_of_elem ? mat->mat_of_elem[e] : 0] == SLD_LAW_SMA160;
    });
  d->color_list.resize(ne);
  {
    std::vector<int32_t> fill(d->color_ptr.begin(), d->color_ptr.end() - 1);
    for (int64_t i = 0; i < ne; ++i) d->color_list[fill[color[d->order[i]]]++] = (int32_t)i;
  }
  auto bail = [&](int rc) {
    sld_dev_destroy(d);
    return rc;
  };
  if (hipGetDevice(&d->device) != hipSuccess) return bail(fail(SMA_ERR_HIP, "no HIP device"));
  if (g_num_cu == 0 &&
      hipDeviceGetAttribute(&g_num_cu, hipDeviceAttributeMultiprocessorCount, d->device) != hipSuccess)
    g_num_cu = 256;
  if (hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(SMA_ERR_HIP, "hipStreamCreate failed"));
  hipStream_t s = d->stream;

  // host staging in processed (colour) order and SoA layouts
  std::vector<int32_t> ln((size_t)ne * 8);
  for (int64_t i = 0; i < ne; ++i)
    std::memcpy(&ln[8 * i], mesh->lnods + 8 * (int64_t)d->order[i], 8 * sizeof(int32_t));
  std::vector<double> soa((size_t)np);
  int rc = 0;
#define TRY(x) do { rc = (x); if (rc) return bail(rc); } while (0)
  TRY(upload(&d->lnods, ln.data(), ln.size(), s)); track(d, d->lnods);
  TRY(upload(&d->color_list_dev, d->color_list.data(), d->color_list.size(), s));
  track(d, d->color_list_dev);
  {
    // slots: the corners at each node in colour order (color_list), so the
    // gather adds them in the order of the coloured scatter.  With at most
    // 16 corners per node the slots are padded to the largest count W
    // (node p owns slots W p .. W p + W - 1)
    int32_t W = 0;
    for (int64_t p = 0; p < np; ++p) W = std::max(W, nadj[p + 1]);
    d->slot_width = (W <= 16 && (int64_t)W * np <= INT32_MAX) ? W : 0;
    for (int64_t p = 0; p < np; ++p) nadj[p + 1] += nadj[p];
    std::vector<int32_t> slot((size_t)ne * 8), fill(nadj.begin(), nadj.end() - 1);
    if (d->slot_width)
      for (int64_t p = 0; p < np; ++p) fill[p] = d->slot_width * (int32_t)p;
    for (int64_t i = 0; i < ne; ++i) {
      const int64_t e = d->color_list[i];
      for (int a = 0; a < 8; ++a) slot[8 * e + a] = fill[ln[8 * e + a]]++;
    }
    TRY(upload(&d->node_ptr, nadj.data(), nadj.size(), s)); track(d, d->node_ptr);
    TRY(upload(&d->corner_slot, slot.data(), slot.size(), s)); track(d, d->corner_slot);
  }
  std::vector<std::vector<double>> stage;
  for (int k = 0; k < 3; ++k) {
    stage.emplace_back((size_t)np);
    for (int64_t p = 0; p < np; ++p) stage.back()[p] = mesh->coord[3 * p + k];
    TRY(upload(&d->coord[k], stage.back().data(), (size_t)np, s)); track(d, d->coord[k]);
    TRY(dalloc(&d->u[k], np)); track(d, d->u[k]);
    TRY(dalloc(&d->v[k], np)); track(d, d->v[k]);
    TRY(dalloc(&d->a[k], np)); track(d, d->a[k]);
    TRY(dalloc(&d->fint64[k], np)); track(d, d->fint64[k]);
    TRY(dalloc(&d->fint32[k], np)); track(d, d->fint32[k]);
    HIPCHK(hipMemsetAsync(d->u[k], 0, np * sizeof(double), s));
    HIPCHK(hipMemsetAsync(d->v[k], 0, np * sizeof(double), s));
    HIPCHK(hipMemsetAsync(d->a[k], 0, np * sizeof(double), s));
    HIPCHK(hipMemsetAsync(d->fint64[k], 0, np * sizeof(double), s));
    HIPCHK(hipMemsetAsync(d->fint32[k], 0, np * sizeof(float), s));
    if (bc->fext0) {
      stage.emplace_back((size_t)np);
      for (int64_t p = 0; p < np; ++p) stage.back()[p] = bc->fext0[3 * p + k];
      TRY(upload(&d->fext0[k], stage.back().data(), (size_t)np, s)); track(d, d->fext0[k]);
    }
  }
  d->fext_func = fext_func;
  {
    // prescribed-dof bits, plus FIX_FORCE where the nodal force pattern is nonzero
    std::vector<uint8_t> fm((size_t)np);
    for (int64_t p = 0; p < np; ++p) {
      fm[p] = (uint8_t)(bc->fixmask[p] & 7);
      if (fext_func >= 0 &&
          (bc->fext0[3 * p] != 0.0 || bc->fext0[3 * p + 1] != 0.0 || bc->fext0[3 * p + 2] != 0.0))
        fm[p] |= (uint8_t)sld::FIX_FORCE;
    }
    TRY(upload(&d->fixmask, fm.data(), (size_t)np, s)); track(d, d->fixmask);
  }
  if (bc->func_of_node) { TRY(upload(&d->func_of_node, bc->func_of_node, (size_t)np, s)); track(d, d->func_of_node); }
  d->nfunc = bc->nfunc;
  d->func_maxpts = bc->func_maxpts;
  if (bc->nfunc > 0) {
    d->func_table_host.assign(bc->func_table, bc->func_table + (size_t)bc->nfunc * bc->func_maxpts * 4);
    d->func_npts_host.assign(bc->func_npts, bc->func_npts + bc->nfunc);
    TRY(upload(&d->func_table, bc->func_table, (size_t)bc->nfunc * bc->func_maxpts * 4, s)); track(d, d->func_table);
    TRY(upload(&d->func_npts, bc->func_npts, (size_t)bc->nfunc, s)); track(d, d->func_npts);
    std::vector<int32_t> sm((size_t)bc->nfunc, 0);
    if (bc->func_smooth) std::memcpy(sm.data(), bc->func_smooth, sm.size() * sizeof(int32_t));
    stage.emplace_back(); // keep staging alive until the stream syncs
    TRY(upload(&d->func_smooth, sm.data(), sm.size(), s)); track(d, d->func_smooth);
    HIPCHK(hipStreamSynchronize(s));
  }
  d->nmat = mat->nmat;
  TRY(upload(&d->props, mat->props, (size_t)mat->nmat * 18, s)); track(d, d->props);
  {
    std::vector<double> th((size_t)mat->nmat * 4);
    for (int k = 0; k < mat->nmat; ++k) sma::mat_thresholds(mat->props + 18 * k, &th[4 * k]);
    TRY(upload(&d->thresholds, th.data(), th.size(), s)); track(d, d->thresholds);
  }
  TRY(upload(&d->rho, mat->rho, (size_t)mat->nmat, s)); track(d, d->rho);
  std::vector<int32_t> me;
  std::vector<double> Ee;
  if (mat->mat_of_elem) {
    me.resize(ne);
    for (int64_t i = 0; i < ne; ++i) me[i] = mat->mat_of_elem[d->order[i]];
    TRY(upload(&d->mat_of_elem, me.data(), me.size(), s)); track(d, d->mat_of_elem);
  }
  if (mat->E_of_elem) {
    Ee.resize(ne);
    for (int64_t i = 0; i < ne; ++i) Ee[i] = mat->E_of_elem[d->order[i]];
    TRY(upload(&d->E_of_elem, Ee.data(), Ee.size(), s)); track(d, d->E_of_elem);
  }
  const int64_t ngp = ne * 8;
  {
    // corner forces (FP64-sized, reused for FP32); padding slots stay zero
    const size_t nslot = d->slot_width ? (size_t)d->slot_width * np : (size_t)ngp;
    TRY(dalloc((double**)&d->efor, nslot * 3)); track(d, d->efor);
    d->efor_bytes = nslot * 3 * sizeof(double);
    HIPCHK(hipMemsetAsync(d->efor, 0, d->efor_bytes, s));
  }
  TRY(dalloc(&d->kg, (size_t)ne * 2)); track(d, d->kg);
  if (any_neo) {
    std::vector<uint8_t> le((size_t)ne);
    for (int64_t i = 0; i < ne; ++i)
      le[i] = mat->law[mat->mat_of_elem ? mat->mat_of_elem[d->order[i]] : 0] == SLD_LAW_NEOHOOKEAN102;
    TRY(upload(&d->law_of_elem, le.data(), le.size(), s)); track(d, d->law_of_elem);
    // stored grouped by law: the neo-Hookean elements are a trailing range
    d->nh_begin = (int32_t)(std::find(le.begin(), le.end(), (uint8_t)1) - le.begin());
  }
  hipLaunchKernelGGL(sld::sld_elem_moduli, dim3(nblocks(ne, 256)), dim3(256), 0, s, ne, d->props,
                     d->mat_of_elem, d->E_of_elem, d->kg);
  for (int sl = 0; sl < 2; ++sl)
    for (int k = 0; k < 3; ++k) {
      TRY(dalloc(&d->sdv[sl][k], ngp)); track(d, d->sdv[sl][k]);
      HIPCHK(hipMemsetAsync(d->sdv[sl][k], 0, ngp * sizeof(double), s));
    }
  TRY(dalloc(&d->counters, 8 * sld::COUNTER_SLOTS)); track(d, d->counters);
  HIPCHK(hipMemsetAsync(d->counters, 0, 8 * sld::COUNTER_SLOTS * sizeof(unsigned long long), s));
  TRY(dalloc(&d->clock, 2)); track(d, d->clock);
  HIPCHK(hipMemsetAsync(d->clock, 0, 2 * sizeof(double), s));
  TRY(dalloc(&d->health, 1)); track(d, d->health);
  HIPCHK(hipStreamSynchronize(s));
  TRY(clear_health(d));
  TRY(dalloc(&d->mass, np)); track(d, d->mass);
  TRY(dalloc(&d->rmass, np)); track(d, d->rmass);
  HIPCHK(hipMemsetAsync(d->mass, 0, np * sizeof(double), s));

  // lumped mass, coloured so the assembly is deterministic
  for (int c = 0; c < ncolor; ++c) {
    sld::MassArgs M{};
    M.e_begin = d->color_ptr[c];
    M.e_end = d->color_ptr[c + 1];
    M.elem_list = d->color_list_dev;
    M.lnods = d->lnods;
    for (int k = 0; k < 3; ++k) M.coord[k] = d->coord[k];
    M.rho = d->rho;
    M.mat_of_elem = d->mat_of_elem;
    M.mass = d->mass;
    const int n = M.e_end - M.e_begin;
    if (n > 0) hipLaunchKernelGGL(sld::sld_mass, dim3(nblocks(n, 64)), dim3(64), 0, s, M);
  }
  hipLaunchKernelGGL(sld::sld_reciprocal, dim3(nblocks(np, 256)), dim3(256), 0, s, d->mass,
                     d->rmass, np);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s));
  d->dtcrit = host_critical_dt(mesh->coord, mesh->lnods, ne, mat->props, mat->rho,
                               mat->mat_of_elem, mat->E_of_elem, mat->law);
#undef TRY
  *out = d;
  return SMA_OK;
}

int sld_dev_set_surrogate(sld_dev* d, sma_mlp* mlp, double tol) {
  if (!d) return fail(SMA_ERR_ARG, "null handle");
  HIPCHK(hipStreamSynchronize(d->stream));
  free_graphs(d);
  if (!mlp) {
    surrogate_detach(d);
    if (d->mode == SLD_MODE_SURROGATE) d->mode = SLD_MODE_FP32;
    return SMA_OK;
  }
  if (!(tol > 0.0)) return fail(SMA_ERR_ARG, "tol must be positive");
  if (!d->sur) {
    auto* S = new sld_dev::Sur();
    d->sur = S;
    const int64_t ngp = d->nelem * 8;
    // the listing pass runs waves of 8 elements (64 points) over the mesh
    S->seg_cap = mlp::list_seg_cap((d->nelem + 7) / 8);
    S->cap = S->seg_cap * mlp::LIST_SEGS;
    const int64_t cap = S->cap, nctr = mlp::LIST_SEGS * mlp::LIST_STRIDE;
    int rc = 0;
    if ((rc = dalloc(&S->status, ngp)) || (rc = dalloc(&S->slot, ngp)) ||
        (rc = dalloc(&S->feat, cap * 16)) || (rc = dalloc(&S->count, nctr)) ||
        (rc = dalloc(&S->Y, cap * 12)) || (rc = dalloc(&S->cand[0], d->nelem)) ||
        (rc = dalloc(&S->cand[1], d->nelem))) {
      surrogate_detach(d);
      return rc;
    }
    HIPCHK(hipMemset(S->count, 0, nctr * sizeof(unsigned)));
    HIPCHK(hipMemset(S->status, 0, ngp));
  }
  if (int rc = surrogate_restart(d)) return rc;
  d->sur->mlp = mlp;
  d->sur->tol = tol;
  return SMA_OK;
}

int sld_dev_set_mode(sld_dev* d, int32_t mode) {
  if (!d) return fail(SMA_ERR_ARG, "null handle");
  if (mode < SLD_MODE_FP64 || mode > SLD_MODE_SURROGATE) return fail(SMA_ERR_ARG, "bad mode");
  if (mode == SLD_MODE_SURROGATE && !surrogate_attached(d))
    return fail(SMA_ERR_STATE, "surrogate mode needs sld_dev_set_surrogate first");
  if (mode != d->mode) {
    HIPCHK(hipStreamSynchronize(d->stream));
    // the corner-force slots change element size (FP64 <-> FP32): clear them
    // so that the padding slots read as zero in the new layout
    if ((mode == SLD_MODE_FP64) != (d->mode == SLD_MODE_FP64)) {
      HIPCHK(hipMemsetAsync(d->efor, 0, d->efor_bytes, d->stream));
      if (d->edamp) HIPCHK(hipMemsetAsync(d->edamp, 0, d->efor_bytes, d->stream));
      HIPCHK(hipStreamSynchronize(d->stream));
    }
    d->mode = mode;
    if (mode == SLD_MODE_SURROGATE)
      if (int rc = surrogate_restart(d)) return rc;
  }
  return SMA_OK;
}

int sld_dev_set_store(sld_dev* d, int32_t store) {
  if (!d) return fail(SMA_ERR_ARG, "null handle");
  if (store && !d->S_out) {
    const int64_t ngp = d->nelem * 8;
    if (dalloc(&d->S_out, 9 * ngp)) return SMA_ERR_HIP;
    track(d, d->S_out);
    if (dalloc(&d->P_out, 9 * ngp)) return SMA_ERR_HIP;
    track(d, d->P_out);
    if (dalloc(&d->flags_out, ngp)) return SMA_ERR_HIP;
    track(d, d->flags_out);
    HIPCHK(hipMemsetAsync(d->S_out, 0, 9 * ngp * sizeof(double), d->stream));
    HIPCHK(hipMemsetAsync(d->P_out, 0, 9 * ngp * sizeof(double), d->stream));
    HIPCHK(hipMemsetAsync(d->flags_out, 0, ngp * sizeof(uint32_t), d->stream));
  }
  d->store = store ? 1 : 0;
  return SMA_OK;
}

int sld_dev_set_tile(sld_dev* d, int32_t epb, int32_t xcd, int32_t persistent) {
  if (!d) return fail(SMA_ERR_ARG, "null handle");
  if (epb != 8 && epb != 16 && epb != 32) return fail(SMA_ERR_ARG, "elems_per_block must be 8, 16 or 32");
  if (persistent < 0 || persistent > 64) return fail(SMA_ERR_ARG, "persistent must be 0..64 workgroups per CU");
  d->epb = epb;
  d->xcd = xcd ? 1 : 0;
  d->persistent = persistent;
  return SMA_OK;
}

int sld_dev_set_assembly(sld_dev* d, int32_t assembly) {
  if (!d) return fail(SMA_ERR_ARG, "null handle");
  if (assembly != SLD_ASSEMBLY_SCATTER && assembly != SLD_ASSEMBLY_GATHER)
    return fail(SMA_ERR_ARG, "assembly must be SLD_ASSEMBLY_SCATTER or SLD_ASSEMBLY_GATHER");
  if (assembly == SLD_ASSEMBLY_SCATTER && d->repeated_nodes)
    return fail(SMA_ERR_STATE, "an element lists a node twice: only the gather assembly is exact");
  if (assembly == SLD_ASSEMBLY_SCATTER && d->damp)
    return fail(SMA_ERR_STATE, "Rayleigh damping runs with the gather assembly only");
  if (assembly != d->assembly) {
    // f_int is complete after every step in both assemblies: nothing to carry
    HIPCHK(hipStreamSynchronize(d->stream));
    d->assembly = assembly;
  }
  return SMA_OK;
}

int sld_dev_set_gravity(sld_dev* d, const double* g) {
  if (!d) return fail(SMA_ERR_ARG, "null handle");
  if (g && !(std::isfinite(g[0]) && std::isfinite(g[1]) && std::isfinite(g[2])))
    return fail(SMA_ERR_ARG, "sld_dev_set_gravity: non-finite g");
  HIPCHK(hipStreamSynchronize(d->stream));
  free_graphs(d);   // captured steps hold the previous value
  d->gravity = g ? 1 : 0;
  for (int k = 0; k < 3; ++k) d->grav[k] = g ? g[k] : 0.0;
  return SMA_OK;
}

int sld_dev_set_body_force(sld_dev* d, int32_t func) {
  if (!d) return fail(SMA_ERR_ARG, "null handle");
  if (func < -1 || func >= d->nfunc)
    return fail(SMA_ERR_ARG, "sld_dev_set_body_force: function index out of range");
  HIPCHK(hipStreamSynchronize(d->stream));
  free_graphs(d);
  d->body_func = func;
  return SMA_OK;
}

int sld_dev_set_damping(sld_dev* d, const double* alpha_beta) {
  if (!d) return fail(SMA_ERR_ARG, "null handle");
  bool any = false;
  if (alpha_beta)
    for (int k = 0; k < 2 * d->nmat; ++k) {
      if (!std::isfinite(alpha_beta[k]) || alpha_beta[k] < 0.0)
        return fail(SMA_ERR_ARG, "sld_dev_set_damping: alpha and beta must be finite and >= 0");
      any |= alpha_beta[k] > 0.0;
    }
  HIPCHK(hipStreamSynchronize(d->stream));
  free_graphs(d);
  if (!any) {
    d->damp = 0;
    return SMA_OK;
  }
  const int64_t np = d->npoin;
  if (!d->damping) {
    int rc = 0;
    if ((rc = dalloc(&d->damping, (size_t)2 * d->nmat))) return rc;
    track(d, d->damping);
    for (int k = 0; k < 3; ++k) {
      if ((rc = dalloc(&d->vdamp[k], np)) || (rc = dalloc(&d->fdamp[k], np))) return rc;
      track(d, d->vdamp[k]);
      track(d, d->fdamp[k]);
      HIPCHK(hipMemset(d->fdamp[k], 0, np * sizeof(double)));
    }
    if ((rc = dalloc((double**)&d->edamp, d->efor_bytes / sizeof(double)))) return rc;
    track(d, d->edamp);
    HIPCHK(hipMemset(d->edamp, 0, d->efor_bytes));
  }
  HIPCHK(hipMemcpy(d->damping, alpha_beta, 2 * d->nmat * sizeof(double), hipMemcpyHostToDevice));
  d->damp = 1;
  d->assembly = SLD_ASSEMBLY_GATHER;
  return SMA_OK;
}

int sld_dev_set_energy(sld_dev* d, int32_t on) {
  if (!d) return fail(SMA_ERR_ARG, "null handle");
  if (on != 0 && on != 1) return fail(SMA_ERR_ARG, "on must be 0 or 1");
  HIPCHK(hipStreamSynchronize(d->stream));
  if (on && !d->energy_acc) {
    const int64_t np = d->npoin;
    int rc = 0;
    for (int k = 0; k < 3 && !rc; ++k) {
      rc = dalloc(&d->du[k], np);
      if (!rc) track(d, d->du[k]);
      if (!rc) rc = dalloc(&d->fint_prev[k], np);
      if (!rc) track(d, d->fint_prev[k]);
      if (!rc) rc = dalloc(&d->fext_prev[k], np);
      if (!rc) track(d, d->fext_prev[k]);
    }
    if (!rc) rc = dalloc(&d->energy_part, 3 * (size_t)nblocks(np, 256));
    if (!rc) track(d, d->energy_part);
    if (!rc) rc = dalloc(&d->energy_acc, 8);
    if (rc) return rc;   // what was allocated is tracked and freed with the handle
    track(d, d->energy_acc);
  }
  if (on) {
    // a fresh balance: no previous internal force, no accumulated work
    const size_t bytes = (size_t)d->npoin * sizeof(double);
    for (int k = 0; k < 3; ++k) {
      HIPCHK(hipMemsetAsync(d->fint_prev[k], 0, bytes, d->stream));
      HIPCHK(hipMemsetAsync(d->fext_prev[k], 0, bytes, d->stream));
    }
    HIPCHK(hipMemsetAsync(d->energy_acc, 0, 8 * sizeof(double), d->stream));
    HIPCHK(hipStreamSynchronize(d->stream));
  }
  d->energy = on;
  return SMA_OK;
}

int sld_dev_set_faces(sld_dev* d, const sld_face_desc* fd) {
  if (!d || !fd) return fail(SMA_ERR_ARG, "null argument");
  const int64_t nf = fd->nface, np = d->npoin;
  if (nf < 0 || nf > (int64_t)INT32_MAX / 4) return fail(SMA_ERR_ARG, "nface out of range");
  if (nf > 0 && (!fd->lnodb || !fd->code || !fd->value))
    return fail(SMA_ERR_ARG, "lnodb, code and value are required");
  if ((fd->transient != 0 && fd->transient != 1) || (fd->follower != 0 && fd->follower != 1))
    return fail(SMA_ERR_ARG, "transient and follower must be 0 or 1");
  for (int64_t f = 0; f < nf; ++f) {
    if (fd->code[f] != 3 && fd->code[f] != 6)
      return fail(SMA_ERR_ARG, "face code must be 3 (pressure) or 6 (traction)");
    for (int a = 0; a < 4; ++a)
      if (fd->lnodb[4 * f + a] < 0 || fd->lnodb[4 * f + a] >= np)
        return fail(SMA_ERR_ARG, "face node out of range");
    if (fd->func && (fd->func[f] < -1 || fd->func[f] >= d->nfunc))
      return fail(SMA_ERR_ARG, "face function out of range");
  }
  // node-sorted corner slots: the corners of a node in ascending face order
  std::vector<int32_t> ptr((size_t)np + 1, 0), slot((size_t)4 * nf);
  for (int64_t k = 0; k < 4 * nf; ++k) ptr[(size_t)fd->lnodb[k] + 1]++;
  for (int64_t p = 0; p < np; ++p) ptr[p + 1] += ptr[p];
  {
    std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
    for (int64_t k = 0; k < 4 * nf; ++k) slot[k] = fill[fd->lnodb[k]]++;
  }
  HIPCHK(hipStreamSynchronize(d->stream));
  free_graphs(d);   // captured steps hold the previous face arguments
  sld_dev::Faces& F = d->faces;
  auto drop = [&](void* q) {
    if (!q) return;
    d->allocs.erase(std::remove(d->allocs.begin(), d->allocs.end(), q), d->allocs.end());
    hipFree(q);
  };
  drop(F.lnodb); drop(F.code); drop(F.value); drop(F.func); drop(F.slot); drop(F.node_ptr);
  drop(F.fface);
  F = sld_dev::Faces{};
  if (nf == 0) return SMA_OK;
  std::vector<int32_t> fn((size_t)nf, -1);
  if (fd->func)
    for (int64_t f = 0; f < nf; ++f) fn[f] = fd->func[f];
  int rc = upload(&F.lnodb, fd->lnodb, (size_t)4 * nf, d->stream);
  if (!rc) { track(d, F.lnodb); rc = upload(&F.code, fd->code, (size_t)nf, d->stream); }
  if (!rc) { track(d, F.code); rc = upload(&F.value, fd->value, (size_t)3 * nf, d->stream); }
  if (!rc) { track(d, F.value); rc = upload(&F.func, fn.data(), (size_t)nf, d->stream); }
  if (!rc) { track(d, F.func); rc = upload(&F.slot, slot.data(), slot.size(), d->stream); }
  if (!rc) { track(d, F.slot); rc = upload(&F.node_ptr, ptr.data(), ptr.size(), d->stream); }
  if (!rc) { track(d, F.node_ptr); rc = dalloc(&F.fface, (size_t)12 * nf); }
  if (rc) {
    F = sld_dev::Faces{};   // tracked buffers are released with the handle
    return rc;
  }
  track(d, F.fface);
  F.nface = nf;
  F.transient = fd->transient;
  F.follower = fd->follower;
  return SMA_OK;
}

int sld_dev_energy(sld_dev* d, double* out) {
  if (!d || !out) return fail(SMA_ERR_ARG, "null argument");
  if (!d->energy_acc) return fail(SMA_ERR_STATE, "the energy balance was never switched on");
  HIPCHK(hipStreamSynchronize(d->stream));
  double acc[5];
  HIPCHK(hipMemcpy(acc, d->energy_acc, sizeof(acc), hipMemcpyDeviceToHost));
  for (int k = 0; k < 5; ++k) out[k] = acc[k];
  return SMA_OK;
}

// The captured run of `len` steps from the current history slot with the
// handle's current settings, built on first use and kept.
static int run_graph(sld_dev* d, double dt, int32_t scheme, double phi, int len, hipGraphExec_t* out) {
  GraphKey key{dt, phi, scheme, d->mode, d->store, d->epb, d->xcd, d->persistent, d->assembly, d->cur,
               d->energy, len, d->damp, d->body_func};
  auto it = d->graphs.find(key);
  if (it == d->graphs.end()) {
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    HIPCHK(hipStreamBeginCapture(d->stream, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    for (int k = 0; k < len && !rc; ++k)
      rc = enqueue_step(d, dt, scheme, phi, (k & 1) ? 1 - d->cur : d->cur, nullptr, k == 0,
                        k + 1 < len);
    hipError_t ec = hipStreamEndCapture(d->stream, &g);
    if (rc) return rc;
    if (ec != hipSuccess) return fail(SMA_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
    HIPCHK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    it = d->graphs.emplace(key, std::make_pair(g, ge)).first;
  }
  *out = it->second.second;
  return SMA_OK;
}

static int step_checks(sld_dev* d, int32_t scheme, double dt, int64_t nsteps) {
  if (!d) return fail(SMA_ERR_ARG, "null handle");
  if (scheme != 0 && scheme != 1) return fail(SMA_ERR_ARG, "scheme must be 0 (CD) or 1 (TW)");
  if (!(dt > 0.0)) return fail(SMA_ERR_ARG, "dt must be positive");
  if (nsteps < 0) return fail(SMA_ERR_ARG, "nsteps must be >= 0");
  if (d->mode == SLD_MODE_SURROGATE && !surrogate_attached(d))
    return fail(SMA_ERR_STATE, "surrogate mode without a surrogate");
  if (d->halted) return fail(SMA_ERR_STATE, d->halt_msg);
  return SMA_OK;
}

int sld_dev_prepare(sld_dev* d, double dt, int32_t scheme, double phi, int64_t nsteps) {
  if (int rc = step_checks(d, scheme, dt, nsteps)) return rc;
  int64_t left = nsteps;
  for (int len : GRAPH_RUNS) {
    if (!d->use_graph || left < len) continue;
    hipGraphExec_t ge = nullptr;
    if (int rc = run_graph(d, dt, scheme, phi, len, &ge)) return rc;
    left %= len;
  }
  return SMA_OK;
}

int sld_dev_step(sld_dev* d, double dt, int32_t scheme, double phi, int64_t nsteps) {
  if (int rc = step_checks(d, scheme, dt, nsteps)) return rc;
  int64_t left = nsteps;
  // runs of 64, 16 and 2 steps replay captured graphs (an even length
  // returns the history and clock slots to where they started); inside a run
  // each corrector also predicts the next step
  for (int len : GRAPH_RUNS) {
    if (!d->use_graph || left < len) continue;
    hipGraphExec_t ge = nullptr;
    if (int rc = run_graph(d, dt, scheme, phi, len, &ge)) return rc;
    for (; left >= len; left -= len) HIPCHK(hipGraphLaunch(ge, d->stream));
  }
  for (; left > 0; --left) {
    int rc = enqueue_step(d, dt, scheme, phi, d->cur);
    if (rc) return rc;
    d->cur = 1 - d->cur;
  }
  return SMA_OK;
}

int sld_dev_time_kernels(sld_dev* d, double dt, int32_t scheme, double phi, int32_t nsteps,
                         double* out) {
  if (!d || !out) return fail(SMA_ERR_ARG, "null argument");
  if (scheme != 0 && scheme != 1) return fail(SMA_ERR_ARG, "scheme must be 0 (CD) or 1 (TW)");
  if (!(dt > 0.0) || nsteps <= 0) return fail(SMA_ERR_ARG, "dt > 0 and nsteps > 0 required");
  if (d->mode == SLD_MODE_SURROGATE && !surrogate_attached(d))
    return fail(SMA_ERR_STATE, "surrogate mode without a surrogate");
  if (d->halted) return fail(SMA_ERR_STATE, d->halt_msg);
  KernelTimer T;
  T.s = d->stream;
  T.elem_reps = 8;
  HIPCHK(hipMalloc((void**)&T.scratch_counters, 8 * sld::COUNTER_SLOTS * sizeof(unsigned long long)));
  struct FreeScratch {
    void* p;
    ~FreeScratch() { (void)hipFree(p); }
  } free_scratch{T.scratch_counters};
  double sum[5] = {0, 0, 0, 0, 0};
  int64_t cnt[5] = {0, 0, 0, 0, 0};
  hipEvent_t t0, t1;
  HIPCHK(hipEventCreate(&t0));
  HIPCHK(hipEventCreate(&t1));
  // all steps are enqueued before the first completes, so each bracket
  // opens on a busy stream; the first one (the first step's predictor)
  // opens on an idle stream and carries its start-up latency: it is left
  // out of the averages when there is more than one step
  HIPCHK(hipEventRecord(t0, d->stream));
  T.used = 0;
  for (int k = 0; k < nsteps; ++k) {
    int rc = enqueue_step(d, dt, scheme, phi, d->cur, &T);
    if (rc) return rc;
    d->cur = 1 - d->cur;
  }
  HIPCHK(hipEventRecord(t1, d->stream));
  HIPCHK(hipEventSynchronize(t1));
  for (size_t i = (nsteps > 1 ? 1 : 0); i < T.used / 2; ++i) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, T.ev[2 * i], T.ev[2 * i + 1]));
    sum[T.kind[i]] += ms;
    cnt[T.kind[i]] += 1;
  }
  float tot = 0.f;
  HIPCHK(hipEventElapsedTime(&tot, t0, t1));
  hipEventDestroy(t0);
  hipEventDestroy(t1);
  if (int rh = read_health(d)) return rh;
  out[0] = cnt[0] ? sum[0] / cnt[0] : 0.0;    // predictor, ms per launch
  // element kernel, ms per launch: the element phase (all its launches,
  // back to back) divided by its number of launches; the gather phase counts
  // as one launch, also where a law-grouped mesh runs it as two kernels
  out[1] = T.element_launches ? sum[1] / T.element_launches : 0.0;
  out[2] = cnt[2] ? sum[2] / cnt[2] : 0.0;    // corrector, ms per launch
  // element launches a step makes (the timing repeats excluded)
  const double per_step = d->assembly == SLD_ASSEMBLY_GATHER
                              ? 1.0 : (double)T.element_launches / nsteps;
  out[3] = per_step;
  // surrogate prepass and inference, ms per launch (0 in the other modes)
  for (int k = 3; k < 5; ++k) out[k + 3] = cnt[k] ? sum[k] / cnt[k] : 0.0;
  out[4] = out[0] + out[2] + out[6] + out[7] + out[1] * per_step;  // kernel ms per step
  out[5] = (double)tot / nsteps;  // eager wall ms per step on the stream, timing repeats included
  return SMA_OK;
}

int sld_dev_autotune(sld_dev* d, double dt, int32_t scheme, double phi, int32_t nsteps,
                     int32_t* best, double* best_ms) {
  if (!d) return fail(SMA_ERR_ARG, "null handle");
  if (scheme != 0 && scheme != 1) return fail(SMA_ERR_ARG, "scheme must be 0 (CD) or 1 (TW)");
  if (!(dt > 0.0) || nsteps <= 0) return fail(SMA_ERR_ARG, "dt > 0 and nsteps > 0 required");
  if (d->mode == SLD_MODE_SURROGATE && !surrogate_attached(d))
    return fail(SMA_ERR_STATE, "surrogate mode without a surrogate");
  hipStream_t s = d->stream;
  HIPCHK(hipStreamSynchronize(s));
  if (int rh = read_health(d)) return rh;
  // snapshot of everything a step writes, restored before every trial and
  // at the end: the state, the assembled and corner forces, the stored
  // stress fields, the energy balance, the counters and the run health
  const int64_t np = d->npoin, ngp = d->nelem * 8;
  std::vector<std::pair<void*, size_t>> live;
  for (int k = 0; k < 3; ++k) {
    live.push_back({d->u[k], np * 8});
    live.push_back({d->v[k], np * 8});
    live.push_back({d->a[k], np * 8});
    live.push_back({d->fint64[k], np * 8});
    live.push_back({d->fint32[k], np * 4});
    for (int sl = 0; sl < 2; ++sl) live.push_back({d->sdv[sl][k], ngp * 8});
    if (d->du[0]) {
      live.push_back({d->du[k], np * 8});
      live.push_back({d->fint_prev[k], np * 8});
      live.push_back({d->fext_prev[k], np * 8});
    }
  }
  live.push_back({d->efor, d->efor_bytes});
  if (d->edamp) {
    live.push_back({d->edamp, d->efor_bytes});
    for (int k = 0; k < 3; ++k) {
      live.push_back({d->vdamp[k], np * 8});
      live.push_back({d->fdamp[k], np * 8});
    }
  }
  if (d->S_out) live.push_back({d->S_out, 9 * ngp * 8});
  if (d->P_out) live.push_back({d->P_out, 9 * ngp * 8});
  if (d->flags_out) live.push_back({d->flags_out, ngp * 4});
  if (d->energy_acc) live.push_back({d->energy_acc, 5 * 8});
  live.push_back({d->clock, 16});
  live.push_back({d->counters, 8 * sld::COUNTER_SLOTS * sizeof(unsigned long long)});
  live.push_back({d->health, sizeof(sld::Health)});
  std::vector<void*> snap(live.size(), nullptr);
  auto cleanup = [&]() { for (void* p : snap) if (p) (void)hipFree(p); };
  for (size_t i = 0; i < live.size(); ++i) {
    if (hipMalloc(&snap[i], live[i].second) != hipSuccess) { cleanup(); return fail(SMA_ERR_HIP, "autotune snapshot alloc"); }
    if (hipMemcpyAsync(snap[i], live[i].first, live[i].second, hipMemcpyDeviceToDevice, s) != hipSuccess) {
      cleanup();
      return fail(SMA_ERR_HIP, "autotune snapshot copy");
    }
  }
  const int cur0 = d->cur;
  auto restore = [&]() -> int {
    for (size_t i = 0; i < live.size(); ++i)
      HIPCHK(hipMemcpyAsync(live[i].first, snap[i], live[i].second, hipMemcpyDeviceToDevice, s));
    d->cur = cur0;
    return 0;
  };
  const int epbs[3] = {8, 16, 32}, pers[3] = {0, 2, 4};
  int b_epb = d->epb, b_xcd = d->xcd, b_per = d->persistent, b_asm = d->assembly;
  double b_ms = 1e300;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  int rc = 0;
  const int asm_first = (d->repeated_nodes || d->damp) ? SLD_ASSEMBLY_GATHER : SLD_ASSEMBLY_SCATTER;
  for (int as = asm_first; as <= SLD_ASSEMBLY_GATHER && !rc; ++as)
  for (int ie = 0; ie < 3 && !rc; ++ie)
    for (int x = 0; x < 2 && !rc; ++x)
      for (int ip = 0; ip < 3 && !rc; ++ip) {
        d->assembly = as;
        d->epb = epbs[ie];
        d->xcd = x;
        d->persistent = pers[ip];
        if ((rc = restore())) break;
        if ((rc = enqueue_step(d, dt, scheme, phi, d->cur))) break;   // warm-up
        d->cur = 1 - d->cur;
        if (hipEventRecord(e0, s) != hipSuccess) { rc = fail(SMA_ERR_HIP, "event"); break; }
        for (int k = 0; k < nsteps && !rc; ++k) {
          rc = enqueue_step(d, dt, scheme, phi, d->cur);
          d->cur = 1 - d->cur;
        }
        if (rc) break;
        float ms = 0.f;
        if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) { rc = fail(SMA_ERR_HIP, "event timing"); break; }
        const double per_step = (double)ms / nsteps;
        if (per_step < b_ms) {
          b_ms = per_step;
          b_epb = d->epb;
          b_xcd = d->xcd;
          b_per = d->persistent;
          b_asm = d->assembly;
        }
      }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  int rr = restore();
  if (!rr) rr = surrogate_restart(d);
  (void)hipStreamSynchronize(s);
  cleanup();
  d->epb = b_epb;
  d->xcd = b_xcd;
  d->persistent = b_per;
  d->assembly = b_asm;
  if (rc) return rc;
  if (rr) return rr;
  if (best) { best[0] = b_epb; best[1] = b_xcd; best[2] = b_per; best[3] = b_asm; }
  if (best_ms) *best_ms = b_ms;
  return SMA_OK;
}

double sld_dev_critical_dt(const sld_dev* d) { return d ? d->dtcrit : -1.0; }

void* sld_dev_stream(sld_dev* d) { return d ? (void*)d->stream : nullptr; }

int sld_dev_sync(sld_dev* d) {
  if (!d) return fail(SMA_ERR_ARG, "null handle");
  HIPCHK(hipStreamSynchronize(d->stream));
  return read_health(d);
}

int sld_dev_coloring(const sld_dev* d, int32_t* ncolor, int32_t* color_ptr, int32_t max_colors,
                     int32_t* elem_order) {
  if (!d || !ncolor) return fail(SMA_ERR_ARG, "null argument");
  *ncolor = d->ncolor;
  if (color_ptr) {
    if (max_colors < d->ncolor) return fail(SMA_ERR_ARG, "color_ptr too small");
    std::memcpy(color_ptr, d->color_ptr.data(), (d->ncolor + 1) * sizeof(int32_t));
  }
  if (elem_order)
    for (int64_t i = 0; i < d->nelem; ++i) elem_order[i] = d->order[d->color_list[i]];
  return SMA_OK;
}

int sld_dev_get(sld_dev* d, int32_t field, void* host, int64_t count) {
  if (!d || !host) return fail(SMA_ERR_ARG, "null argument");
  HIPCHK(hipStreamSynchronize(d->stream));
  const int64_t np = d->npoin, ne = d->nelem, ngp = ne * 8;
  // the fields stay readable after a halt (the state the failure left);
  // the host copy of the halt is refreshed so the next step refuses to run
  sld::Health h{};
  {
    const std::string keep = sma_abi::last_error;
    (void)read_health(d, &h);
    sma_abi::last_error = keep;
  }
  auto need = [&](int64_t n) { return count == n ? 0 : fail(SMA_ERR_ARG, "count does not match field size"); };
  std::vector<double> tmp;
  switch (field) {
    case SLD_FIELD_U: case SLD_FIELD_V: case SLD_FIELD_A: case SLD_FIELD_FINT: {
      if (need(3 * np)) return SMA_ERR_ARG;
      double* out = (double*)host;
      tmp.resize(np);
      for (int k = 0; k < 3; ++k) {
        if (field == SLD_FIELD_FINT && d->mode != SLD_MODE_FP64) {
          std::vector<float> t32(np);
          HIPCHK(hipMemcpy(t32.data(), d->fint32[k], np * sizeof(float), hipMemcpyDeviceToHost));
          for (int64_t p = 0; p < np; ++p) tmp[p] = t32[p];
        } else {
          double* src = field == SLD_FIELD_U ? d->u[k] : field == SLD_FIELD_V ? d->v[k]
                      : field == SLD_FIELD_A ? d->a[k] : d->fint64[k];
          HIPCHK(hipMemcpy(tmp.data(), src, np * sizeof(double), hipMemcpyDeviceToHost));
        }
        for (int64_t p = 0; p < np; ++p) out[3 * p + k] = tmp[p];
      }
      return SMA_OK;
    }
    case SLD_FIELD_MASS:
      if (need(np)) return SMA_ERR_ARG;
      HIPCHK(hipMemcpy(host, d->mass, np * sizeof(double), hipMemcpyDeviceToHost));
      return SMA_OK;
    case SLD_FIELD_SDV: {
      if (need(3 * ngp)) return SMA_ERR_ARG;
      double* out = (double*)host;
      const int k_slot = d->cur;  // after a step the slot just written holds TIME_N = ITER_K
      tmp.resize(ngp);
      for (int k = 0; k < 3; ++k) {
        HIPCHK(hipMemcpy(tmp.data(), d->sdv[k_slot][k], ngp * sizeof(double), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < ne; ++i)
          for (int g = 0; g < 8; ++g) out[(8 * (int64_t)d->order[i] + g) * 3 + k] = tmp[8 * i + g];
      }
      return SMA_OK;
    }
    case SLD_FIELD_REACTIONS: {
      if (need(3 * np)) return SMA_ERR_ARG;
      std::vector<double> f(3 * np);
      int rc = sld_dev_get(d, SLD_FIELD_FINT, f.data(), 3 * np);
      if (rc) return rc;
      std::vector<uint8_t> fix(np);
      HIPCHK(hipMemcpy(fix.data(), d->fixmask, np, hipMemcpyDeviceToHost));
      // frxid = elrhs = elfext - elfint - the damping force on prescribed
      // dofs: the element loads are m (g + b(t)) of gravity and the body
      // force (the Neumann and nodal loads are assembled apart)
      const bool load = d->gravity || d->body_func >= 0;
      std::vector<double> m(load ? np : 0), fd(d->damp ? 3 * np : 0);
      if (load) HIPCHK(hipMemcpy(m.data(), d->mass, np * sizeof(double), hipMemcpyDeviceToHost));
      if (d->damp)
        for (int k = 0; k < 3; ++k) {
          std::vector<double> t(np);
          HIPCHK(hipMemcpy(t.data(), d->fdamp[k], np * sizeof(double), hipMemcpyDeviceToHost));
          for (int64_t p = 0; p < np; ++p) fd[3 * p + k] = t[p];
        }
      double gb[3] = {0.0, 0.0, 0.0};
      if (load) {
        double t = 0.0;
        HIPCHK(hipMemcpy(&t, d->clock + d->cur, sizeof(double), hipMemcpyDeviceToHost));
        for (int k = 0; k < 3; ++k) {
          gb[k] = d->gravity ? d->grav[k] : 0.0;
          if (d->body_func >= 0)
            gb[k] = gb[k] + host_funbou_body(&d->func_table_host[(size_t)d->body_func * d->func_maxpts * 4],
                                             d->func_npts_host[d->body_func], k, t);
        }
      }
      double* out = (double*)host;
      for (int64_t p = 0; p < np; ++p)
        for (int k = 0; k < 3; ++k) {
          double r = 0.0 - f[3 * p + k];
          if (d->damp) r = r - fd[3 * p + k];
          if (load) r = r + m[p] * gb[k];
          out[3 * p + k] = ((fix[p] >> k) & 1) ? r : 0.0;
        }
      return SMA_OK;
    }
    case SLD_FIELD_STRESS: case SLD_FIELD_PK1: {
      if (!d->S_out) return fail(SMA_ERR_STATE, "stress not stored (sld_dev_set_store)");
      if (need(9 * ngp)) return SMA_ERR_ARG;
      double* out = (double*)host;
      tmp.resize(9 * ngp);
      HIPCHK(hipMemcpy(tmp.data(), field == SLD_FIELD_PK1 ? d->P_out : d->S_out,
                       9 * ngp * sizeof(double), hipMemcpyDeviceToHost));
      for (int k = 0; k < 9; ++k)
        for (int64_t i = 0; i < ne; ++i)
          for (int g = 0; g < 8; ++g)
            out[(8 * (int64_t)d->order[i] + g) * 9 + k] = tmp[k * ngp + 8 * i + g];
      return SMA_OK;
    }
    case SLD_FIELD_FLAGS: {
      if (!d->flags_out) return fail(SMA_ERR_STATE, "flags not stored (sld_dev_set_store)");
      if (need(ngp)) return SMA_ERR_ARG;
      std::vector<uint32_t> t(ngp);
      HIPCHK(hipMemcpy(t.data(), d->flags_out, ngp * sizeof(uint32_t), hipMemcpyDeviceToHost));
      uint32_t* out = (uint32_t*)host;
      for (int64_t i = 0; i < ne; ++i)
        for (int g = 0; g < 8; ++g) out[8 * (int64_t)d->order[i] + g] = t[8 * i + g];
      return SMA_OK;
    }
    case SLD_FIELD_TIME:
      if (need(1)) return SMA_ERR_ARG;
      HIPCHK(hipMemcpy(host, d->clock + d->cur, sizeof(double), hipMemcpyDeviceToHost));
      return SMA_OK;
    case SLD_FIELD_HEALTH: {
      if (need(4)) return SMA_ERR_ARG;
      double* out = (double*)host;
      out[0] = (double)h.halted;
      out[1] = -1.0;
      out[2] = 0.0;
      out[3] = -1.0;
      if (h.halted & sld::HALT_DETF) {
        const int64_t e = (int64_t)(0xffffffffu - (uint32_t)(h.detf & 0xffffffffull));
        const uint32_t bits = (uint32_t)(h.detf >> 32);
        float mag;
        std::memcpy(&mag, &bits, sizeof(mag));
        out[1] = (double)((e >= 0 && e < ne) ? d->order[e] : e);
        out[2] = -(double)mag;
      }
      if (h.halted & sld::HALT_NONFINITE) out[3] = (double)h.node;
      return SMA_OK;
    }
    case SLD_FIELD_COUNTERS: {
      if (need(8)) return SMA_ERR_ARG;
      std::vector<unsigned long long> slots(8 * sld::COUNTER_SLOTS);
      HIPCHK(hipMemcpy(slots.data(), d->counters, slots.size() * sizeof(unsigned long long),
                       hipMemcpyDeviceToHost));
      unsigned long long* out = (unsigned long long*)host;
      for (int k = 0; k < 8; ++k) out[k] = 0;
      for (int b = 0; b < sld::COUNTER_SLOTS; ++b)
        for (int k = 0; k < 8; ++k) out[k] += slots[8 * b + k];
      return SMA_OK;
    }
    default:
      return fail(SMA_ERR_ARG, "unknown field");
  }
}

int sld_dev_set(sld_dev* d, int32_t field, const void* host, int64_t count) {
  if (!d || !host) return fail(SMA_ERR_ARG, "null argument");
  HIPCHK(hipStreamSynchronize(d->stream));
  const int64_t np = d->npoin, ne = d->nelem, ngp = ne * 8;
  std::vector<double> tmp;
  // a state set by the caller (restart) starts a healthy run
  if (field == SLD_FIELD_U || field == SLD_FIELD_V || field == SLD_FIELD_A ||
      field == SLD_FIELD_SDV || field == SLD_FIELD_TIME) {
    if (int rc = clear_health(d)) return rc;
    if (int rc = surrogate_restart(d)) return rc;
  }
  switch (field) {
    case SLD_FIELD_U: case SLD_FIELD_V: case SLD_FIELD_A: {
      if (count != 3 * np) return fail(SMA_ERR_ARG, "count does not match field size");
      const double* in = (const double*)host;
      tmp.resize(np);
      for (int k = 0; k < 3; ++k) {
        for (int64_t p = 0; p < np; ++p) tmp[p] = in[3 * p + k];
        double* dst = field == SLD_FIELD_U ? d->u[k] : field == SLD_FIELD_V ? d->v[k] : d->a[k];
        HIPCHK(hipMemcpy(dst, tmp.data(), np * sizeof(double), hipMemcpyHostToDevice));
      }
      return SMA_OK;
    }
    case SLD_FIELD_SDV: {
      if (count != 3 * ngp) return fail(SMA_ERR_ARG, "count does not match field size");
      const double* in = (const double*)host;
      tmp.resize(ngp);
      for (int k = 0; k < 3; ++k) {
        for (int64_t i = 0; i < ne; ++i)
          for (int g = 0; g < 8; ++g) tmp[8 * i + g] = in[(8 * (int64_t)d->order[i] + g) * 3 + k];
        for (int sl = 0; sl < 2; ++sl)
          HIPCHK(hipMemcpy(d->sdv[sl][k], tmp.data(), ngp * sizeof(double), hipMemcpyHostToDevice));
      }
      return SMA_OK;
    }
    case SLD_FIELD_TIME:
      if (count != 1) return fail(SMA_ERR_ARG, "count must be 1");
      HIPCHK(hipMemcpy(d->clock + d->cur, host, sizeof(double), hipMemcpyHostToDevice));
      return SMA_OK;
    default:
      return fail(SMA_ERR_ARG, "field cannot be set");
  }
}

int sma_device_malloc(void** p, int64_t bytes) {
  if (!p || bytes < 0) return fail(SMA_ERR_ARG, "sma_device_malloc: bad arguments");
  *p = nullptr;
  HIPCHK(hipMalloc(p, bytes > 0 ? (size_t)bytes : 1));
  return SMA_OK;
}

int sma_device_free(void* p) {
  if (p) HIPCHK(hipFree(p));
  return SMA_OK;
}

int sma_memcpy(void* dst, const void* src, int64_t bytes, int32_t kind) {
  if (bytes < 0 || (bytes > 0 && (!dst || !src))) return fail(SMA_ERR_ARG, "sma_memcpy: bad arguments");
  if (bytes == 0) return SMA_OK;
  const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost
                                                             : hipMemcpyDeviceToDevice;
  HIPCHK(hipMemcpy(dst, src, (size_t)bytes, k));
  return SMA_OK;
}

// A plain streaming copy (16 B per lane per iteration), the reference point
// for HBM byte counters of the other kernels.
__global__ void sma_copy_kernel(const double2* __restrict__ src, double2* __restrict__ dst, int64_t n2) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

int sma_copy_probe(const void* src, void* dst, int64_t bytes, void* stream) {
  if (!src || !dst || bytes <= 0 || bytes % 16) return fail(SMA_ERR_ARG, "sma_copy_probe: bytes must be a positive multiple of 16");
  const int64_t n2 = bytes / 16;
  hipLaunchKernelGGL(sma_copy_kernel, dim3(2048), dim3(256), 0, (hipStream_t)stream,
                     (const double2*)src, (double2*)dst, n2);
  HIPCHK(hipGetLastError());
  return SMA_OK;
}

int sma_device_synchronize(void) {
  HIPCHK(hipDeviceSynchronize());
  return SMA_OK;
}

int sma_device_count(int32_t* n) {
  if (!n) return fail(SMA_ERR_ARG, "null argument");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  *n = (e == hipSuccess) ? c : 0;
  return SMA_OK;
}

int sld_dev_destroy(sld_dev* d) {
  if (!d) return SMA_OK;
  if (d->stream) hipStreamSynchronize(d->stream);
  free_graphs(d);
  surrogate_detach(d);
  for (void* p : d->allocs) hipFree(p);
  if (d->stream) hipStreamDestroy(d->stream);
  delete d;
  return SMA_OK;
}

// Staging of sma160_eval, one per host thread: a device block and a pinned
// host block, grown on demand and kept for the thread's life, so a call
// (the reference evaluates the law once per element, 8 points) costs one
// upload, one launch and one download instead of allocations per array.
struct EvalStage {
  void* dev = nullptr;
  void* host = nullptr;
  size_t bytes = 0;
  int device = -1;               // the device that holds `dev`
  int reserve(size_t want) {
    int cur = 0;
    HIPCHK(hipGetDevice(&cur));
    if (cur != device) {
      // another device is current: a block of its own (the other device's
      // block stays with it until the process ends)
      dev = nullptr;
      bytes = 0;
      device = cur;
    }
    if (want <= bytes) return SMA_OK;
    size_t nb = bytes ? bytes : 4096;
    while (nb < want) nb *= 2;
    if (dev) (void)hipFree(dev);
    if (host) (void)hipHostFree(host);
    dev = host = nullptr;
    bytes = 0;
    HIPCHK(hipMalloc(&dev, nb));
    HIPCHK(hipHostMalloc(&host, nb, hipHostMallocDefault));
    bytes = nb;
    return SMA_OK;
  }
};
// the blocks are left to the process's end (no destructor: freeing after
// the runtime has shut down is not safe)
EvalStage& eval_stage() {
  static thread_local EvalStage* st = new EvalStage();
  return *st;
}

static int law_launch(int64_t n, int32_t nmat, const double* props, const int32_t* mat, const double* E,
                      const double* F, const double* sdv_old, double* sdv_new, double* S,
                      uint32_t* flags, int32_t precision, hipStream_t s) {
  sld::LawArgs A{n, nmat, props, mat, E, F, sdv_old, sdv_new, S, flags};
  if (precision == 0)
    hipLaunchKernelGGL((sld::sma160_points<false>), dim3(nblocks(n, 256)), dim3(256), 0, s, A);
  else
    hipLaunchKernelGGL((sld::sma160_points<true>), dim3(nblocks(n, 256)), dim3(256), 0, s, A);
  HIPCHK(hipGetLastError());
  return SMA_OK;
}

int sma160_eval_device(int64_t n, int32_t n_mat, const double* props, const int32_t* mat_of_gp,
                       const double* E_of_gp, const double* F, const double* T,
                       const double* sdv_old, double* sdv_new, double* S, uint32_t* flags,
                       int32_t precision, void* stream) {
  (void)T;
  if (n < 0 || n_mat <= 0 || (n > 0 && (!props || !F || !sdv_old || !sdv_new || !S)))
    return fail(SMA_ERR_ARG, "sma160_eval_device: bad arguments");
  if (precision != 0 && precision != 1) return fail(SMA_ERR_ARG, "precision must be 0 or 1");
  if (n == 0) return SMA_OK;
  return law_launch(n, n_mat, props, mat_of_gp, E_of_gp, F, sdv_old, sdv_new, S, flags, precision,
                    (hipStream_t)stream);
}

int sma160_eval(int64_t n, int32_t n_mat, const double* props, const int32_t* mat_of_gp,
                const double* E_of_gp, const double* F, const double* T, const double* sdv_old,
                double* sdv_new, double* S, uint32_t* flags, int32_t precision) {
  (void)T;
  if (n < 0 || n_mat <= 0 || (n > 0 && (!props || !F || !sdv_old || !sdv_new || !S)))
    return fail(SMA_ERR_ARG, "sma160_eval: bad arguments");
  if (precision != 0 && precision != 1) return fail(SMA_ERR_ARG, "precision must be 0 or 1");
  if (n == 0) return SMA_OK;
  if (mat_of_gp)
    for (int64_t i = 0; i < n; ++i)
      if (mat_of_gp[i] < 0 || mat_of_gp[i] >= n_mat)
        return fail(SMA_ERR_ARG, "sma160_eval: material index out of range");
  // inputs, then outputs, each section 8-byte aligned, in one block
  const size_t b_props = (size_t)n_mat * 18 * 8, b_mat = mat_of_gp ? (size_t)((n * 4 + 7) / 8 * 8) : 0,
               b_E = E_of_gp ? (size_t)n * 8 : 0, b_F = (size_t)n * 72, b_so = (size_t)n * 24;
  const size_t b_sn = (size_t)n * 24, b_S = (size_t)n * 72, b_fl = flags ? (size_t)((n * 4 + 7) / 8 * 8) : 0;
  const size_t in_bytes = b_props + b_mat + b_E + b_F + b_so, out_bytes = b_sn + b_S + b_fl;
  EvalStage& st = eval_stage();
  if (int rc = st.reserve(in_bytes + out_bytes)) return rc;
  char* h = static_cast<char*>(st.host);
  char* dv = static_cast<char*>(st.dev);
  size_t o = 0;
  auto put = [&](const void* src, size_t bytes, size_t span) -> size_t {
    if (src) std::memcpy(h + o, src, bytes);
    const size_t at = o;
    o += span;
    return at;
  };
  const size_t o_p = put(props, b_props, b_props), o_m = put(mat_of_gp, (size_t)n * 4, b_mat);
  const size_t o_E = put(E_of_gp, b_E, b_E), o_F = put(F, b_F, b_F), o_so = put(sdv_old, b_so, b_so);
  const size_t o_sn = in_bytes, o_S = o_sn + b_sn, o_fl = o_S + b_S;
  HIPCHK(hipMemcpy(dv, h, in_bytes, hipMemcpyHostToDevice));
  auto D = [&](size_t off, size_t bytes) -> void* { return bytes ? dv + off : nullptr; };
  if (int rc = law_launch(n, n_mat, (const double*)D(o_p, b_props), (const int32_t*)D(o_m, b_mat),
                          (const double*)D(o_E, b_E), (const double*)D(o_F, b_F),
                          (const double*)D(o_so, b_so), (double*)D(o_sn, b_sn), (double*)D(o_S, b_S),
                          (uint32_t*)D(o_fl, b_fl), precision, nullptr))
    return rc;
  HIPCHK(hipMemcpy(h + in_bytes, dv + in_bytes, out_bytes, hipMemcpyDeviceToHost));
  std::memcpy(sdv_new, h + o_sn, b_sn);
  std::memcpy(S, h + o_S, b_S);
  if (flags) std::memcpy(flags, h + o_fl, (size_t)n * 4);
  return SMA_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// sma160_eval without device allocation.  sma160_eval stages the caller's
// host arrays through device buffers it allocates and frees on every call;
// sma160_eval_noalloc stages them through `scratch`, a fixed array in the
// library's device memory.  Being a file-scope __device__ object, every GPU
// the library runs on holds its own copy (at its own address) until the
// process ends.  Host code never uses the array's name as an address: the
// first call on a device looks that device's copy up with
// hipGetSymbolAddress, with the device current, and keeps the pointer per
// device.  Every call on a GPU shares that GPU's copy and leaves its
// contents behind (the law kernel reads only what the call copied in, so
// nothing needs clearing); calls that overlap on one GPU race on it.
// ---------------------------------------------------------------------------
namespace {

constexpr int64_t kNoallocMaxGp = SMA160_NOALLOC_MAX_GP;
constexpr int32_t kNoallocMaxMat = SMA160_NOALLOC_MAX_MAT;
constexpr int kNoallocMaxDevices = 64;
// staging layout in doubles: props [max mat][18], then per point E (1), F (9),
// history in (3), history out (3), S (9), and the material index and flags as
// two 32-bit words (one double's room)
constexpr int64_t kNoallocPropsDoubles = (int64_t)kNoallocMaxMat * 18;
constexpr int64_t kNoallocPerGpDoubles = 1 + 9 + 3 + 3 + 9 + 1;

__device__ __align__(256) double scratch[kNoallocPropsDoubles + kNoallocPerGpDoubles * kNoallocMaxGp];

std::mutex g_noalloc_mu;
double* g_noalloc_ptr[kNoallocMaxDevices] = {};

// Device address of the current device's copy of `scratch`.
int noalloc_scratch(double** out) {
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoallocMaxDevices)
    return fail(SMA_ERR_HIP, "sma160_eval_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(SMA_ERR_HIP, "sma160_eval_noalloc: no device address for the staging array");
    g_noalloc_ptr[dev] = (double*)p;
  }
  *out = g_noalloc_ptr[dev];
  return SMA_OK;
}

}  // namespace

extern "C" int sma160_eval_noalloc(int64_t n, int32_t n_mat, const double* props,
                                   const int32_t* mat_of_gp, const double* E_of_gp,
                                   const double* F, const double* T, const double* sdv_old,
                                   double* sdv_new, double* S, uint32_t* flags,
                                   int32_t precision) {
  (void)T;
  if (n < 0 || n_mat <= 0 || (n > 0 && (!props || !F || !sdv_old || !sdv_new || !S)))
    return fail(SMA_ERR_ARG, "sma160_eval_noalloc: bad arguments");
  if (precision != 0 && precision != 1) return fail(SMA_ERR_ARG, "precision must be 0 or 1");
  if (n > kNoallocMaxGp)
    return fail(SMA_ERR_ARG, "sma160_eval_noalloc: n_gp = " + std::to_string(n) +
                                 " exceeds the maximum " + std::to_string(kNoallocMaxGp) +
                                 " (SMA160_NOALLOC_MAX_GP)");
  if (n_mat > kNoallocMaxMat)
    return fail(SMA_ERR_ARG, "sma160_eval_noalloc: n_mat = " + std::to_string(n_mat) +
                                 " exceeds the maximum " + std::to_string(kNoallocMaxMat) +
                                 " (SMA160_NOALLOC_MAX_MAT)");
  if (n == 0) return SMA_OK;
  if (mat_of_gp)
    for (int64_t i = 0; i < n; ++i)
      if (mat_of_gp[i] < 0 || mat_of_gp[i] >= n_mat)
        return fail(SMA_ERR_ARG, "sma160_eval_noalloc: material index out of range");
  double* base = nullptr;
  if (int rc = noalloc_scratch(&base)) return rc;
  // this call's arrays, sized by n (the kernel's SoA stride is n)
  double* d_props = base;
  double* d_E = d_props + kNoallocPropsDoubles;
  double* d_F = d_E + n;
  double* d_sold = d_F + 9 * n;
  double* d_snew = d_sold + 3 * n;
  double* d_S = d_snew + 3 * n;
  int32_t* d_mat = (int32_t*)(d_S + 9 * n);
  uint32_t* d_flags = (uint32_t*)(d_mat + n);
  HIPCHK(hipMemcpy(d_props, props, (size_t)n_mat * 18 * sizeof(double), hipMemcpyHostToDevice));
  if (E_of_gp) HIPCHK(hipMemcpy(d_E, E_of_gp, (size_t)n * sizeof(double), hipMemcpyHostToDevice));
  if (mat_of_gp) HIPCHK(hipMemcpy(d_mat, mat_of_gp, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d_F, F, (size_t)9 * n * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d_sold, sdv_old, (size_t)3 * n * sizeof(double), hipMemcpyHostToDevice));
  // the null stream, as sma160_eval
  if (int rc = law_launch(n, n_mat, d_props, mat_of_gp ? d_mat : nullptr, E_of_gp ? d_E : nullptr,
                          d_F, d_sold, d_snew, d_S, flags ? d_flags : nullptr, precision, nullptr))
    return rc;
  HIPCHK(hipMemcpy(sdv_new, d_snew, (size_t)3 * n * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(S, d_S, (size_t)9 * n * sizeof(double), hipMemcpyDeviceToHost));
  if (flags) HIPCHK(hipMemcpy(flags, d_flags, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost));
  return SMA_OK;
}
