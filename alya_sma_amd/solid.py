"""Host-side mirror of the reference's solidz operator surface for the
explicit SMA step, over the C ABI of include/alya_sma.h.

  eval_sld_stress_model_160(...)   -> sm160_eval / sm160_eval_device
      (sld_stress_model_160.f90:36-48, explicit call with flagt = 0)
  Solidz.step(...)                 -> sld_dev_step
      (sld_explicit / sld_explicit_tw, mod_sld_solution_methods.f90:101, :385)
  Solidz.get(...) / .set(...)      -> sld_dev_get / sld_dev_set
      (displ, veloc_sld, accel_sld, finte_sld, vmass_sld, svegm_sld)

Only plumbing lives here: array marshalling and error translation.  Every
computation runs in libalya_sma_amd.so on the GPU.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _c(a, dt):
    return None if a is None else np.ascontiguousarray(a, dtype=dt)


def eval_sld_stress_model_160(props, F, sdv_old, mat_of_gp=None, E_of_gp=None, precision=0):
    """Point-wise SMA law on the GPU (host arrays in, host arrays out).

    props [nmat, 18]; F [9, n] column-major 3x3 per point; sdv_old [3, n]
    = (xS, F_AS, F_SA) at TIME_N.  Returns (sdv_new [3, n], S [9, n], flags [n]).
    """
    props = _c(np.atleast_2d(props), np.float64)
    F = _c(F, np.float64)
    sdv_old = _c(sdv_old, np.float64)
    n = F.shape[1]
    if F.shape != (9, n) or sdv_old.shape != (3, n):
        raise ValueError("F must be [9, n] and sdv_old [3, n]")
    sdv_new = np.zeros((3, n))
    S = np.zeros((9, n))
    flags = np.zeros(n, dtype=np.uint32)
    mat = _c(mat_of_gp, np.int32)
    E = _c(E_of_gp, np.float64)
    check(lib().sma160_eval(n, props.shape[0], _p(props), _p(mat), _p(E), _p(F), None,
                            _p(sdv_old), _p(sdv_new), _p(S), _p(flags), int(precision)),
          "sma160_eval")
    return sdv_new, S, flags


def eval_sld_stress_model_160_device(n, nmat, props_ptr, F_ptr, sdv_old_ptr, sdv_new_ptr,
                                     S_ptr, flags_ptr=None, mat_ptr=None, E_ptr=None,
                                     precision=0, stream=None):
    """Same law on device pointers (e.g. torch tensors' data_ptr()), async."""
    check(lib().sma160_eval_device(n, nmat, props_ptr, mat_ptr, E_ptr, F_ptr, None, sdv_old_ptr,
                                   sdv_new_ptr, S_ptr, flags_ptr, int(precision), stream),
          "sma160_eval_device")


class Solidz:
    """Device state of one explicit SMA problem (see alya_sma_amd.problems)."""

    def __init__(self, prob):
        L = lib()
        self._keep = []
        k = self._keep.append

        def arr(x, dt):
            if x is None:
                return None
            a = np.ascontiguousarray(x, dtype=dt)
            k(a)
            return a

        coord = arr(prob["coord"], np.float64)
        lnods = arr(prob["lnods"], np.int32)
        self.npoin, self.nelem = coord.shape[0], lnods.shape[0]
        mesh = _lib.SldMeshDesc(self.npoin, self.nelem, _p(lnods), _p(coord))
        props = arr(np.atleast_2d(prob["props"]), np.float64)
        matd = _lib.SldMaterialDesc(props.shape[0], _p(props), _p(arr(prob["rho"], np.float64)),
                                    _p(arr(prob.get("mat_of_elem"), np.int32)),
                                    _p(arr(prob.get("E_of_elem"), np.float64)),
                                    _p(arr(prob.get("law_of_mat"), np.int32)))
        ftab = arr(prob["func_table"], np.float64)
        bc = _lib.SldBcDesc(_p(arr(prob["fixmask"], np.uint8)),
                            _p(arr(prob.get("func_of_node"), np.int32)),
                            ftab.shape[0], ftab.shape[1],
                            _p(arr(prob["func_npts"], np.int32)),
                            _p(arr(prob["func_smooth"], np.int32)), _p(ftab),
                            _p(arr(prob.get("fext0"), np.float64)),
                            int(prob.get("fext_func", -1)))
        h = C.c_void_p()
        check(L.sld_dev_create(C.byref(mesh), C.byref(matd), C.byref(bc), C.byref(h)),
              "sld_dev_create")
        self.h = h
        self._keep = []
        if prob.get("faces") is not None:
            self.set_faces(prob["faces"])
        if prob.get("gravity") is not None:
            self.set_gravity(prob["gravity"])
        if prob.get("damping") is not None:
            self.set_damping(prob["damping"])
        if prob.get("body_func", -1) >= 0:
            self.set_body_force(prob["body_func"])

    def close(self):
        if getattr(self, "h", None):
            lib().sld_dev_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---------------------------------------------------------------- control
    def set_mode(self, mode):
        check(lib().sld_dev_set_mode(self.h, int(mode)), "sld_dev_set_mode")

    def set_store(self, on=True):
        check(lib().sld_dev_set_store(self.h, 1 if on else 0), "sld_dev_set_store")

    def set_tile(self, elems_per_block, xcd_remap=True, persistent=0):
        check(lib().sld_dev_set_tile(self.h, int(elems_per_block), 1 if xcd_remap else 0,
                                     int(persistent)), "sld_dev_set_tile")

    def set_gravity(self, g):
        """Gravity acceleration g[3] (grnor_sld * gravi_sld): f_ext += m g at
        every node; None removes it."""
        if g is None:
            check(lib().sld_dev_set_gravity(self.h, None), "sld_dev_set_gravity")
            return
        gv = (C.c_double * 3)(*[float(x) for x in g])
        check(lib().sld_dev_set_gravity(self.h, gv), "sld_dev_set_gravity")

    def set_damping(self, alpha_beta):
        """Rayleigh damping (alpha, beta) per material, [nmat, 2]
        (RAYLEIGH_DAMPING: RAYALPHA, RAYBETA); None removes it."""
        if alpha_beta is None:
            check(lib().sld_dev_set_damping(self.h, None), "sld_dev_set_damping")
            return
        ab = np.ascontiguousarray(np.atleast_2d(alpha_beta), dtype=np.float64)
        check(lib().sld_dev_set_damping(self.h, _p(ab)), "sld_dev_set_damping")

    def set_body_force(self, func):
        """Body force from load function `func` (BODYF); -1 removes it."""
        check(lib().sld_dev_set_body_force(self.h, int(func)), "sld_dev_set_body_force")

    def set_surrogate(self, surrogate, tol=1e-4):
        """Attach a trained alya_sma_amd.surrogate.Surrogate: mode 2 then
        takes its xS update on single-branch transforming points that lie
        within tol (in xS) of the return map's solution; None detaches."""
        h = surrogate.h if surrogate is not None else None
        check(lib().sld_dev_set_surrogate(self.h, h, float(tol)), "sld_dev_set_surrogate")

    def set_assembly(self, assembly):
        """"scatter" (one element launch per colour) or "gather" (one launch,
        per-node sum in the corrector); both give bitwise equal results."""
        a = {"scatter": 0, "gather": 1}.get(assembly, assembly)
        check(lib().sld_dev_set_assembly(self.h, int(a)), "sld_dev_set_assembly")

    def set_faces(self, faces):
        """Neumann loads on boundary faces: dict(lnodb [nf, 4], code [nf]
        (3 pressure, 6 traction), value [nf, 3], func [nf] or None,
        transient, follower); None or nf = 0 removes them."""
        keep = []

        def arr(x, dt):
            if x is None:
                return None
            a = np.ascontiguousarray(x, dtype=dt)
            keep.append(a)
            return a

        if faces is None or len(faces["code"]) == 0:
            desc = _lib.SldFaceDesc(0, None, None, None, None, 0, 0)
        else:
            desc = _lib.SldFaceDesc(len(faces["code"]), _p(arr(faces["lnodb"], np.int32)),
                                    _p(arr(faces["code"], np.int32)),
                                    _p(arr(faces["value"], np.float64)),
                                    _p(arr(faces.get("func"), np.int32)),
                                    int(bool(faces.get("transient", False))),
                                    int(bool(faces.get("follower", False))))
        check(lib().sld_dev_set_faces(self.h, C.byref(desc)), "sld_dev_set_faces")

    def set_energy(self, on=True):
        """Start (True) a fresh energy balance or stop (False) accumulating."""
        check(lib().sld_dev_set_energy(self.h, 1 if on else 0), "sld_dev_set_energy")

    def energy(self):
        """{"allie", "allwk", "allke", "etota", "steps"} of the running balance
        (mod_sld_energy.f90:141-162)."""
        out = np.zeros(5, np.float64)
        check(lib().sld_dev_energy(self.h, out.ctypes.data), "sld_dev_energy")
        return dict(zip(("allie", "allwk", "allke", "etota", "steps"), out.tolist()))

    def autotune(self, dt, nsteps=3, scheme=0, phi=1.033):
        """Pick the fastest assembly and element-kernel launch shape; state
        is unchanged."""
        best = np.zeros(4, dtype=np.int32)
        ms = C.c_double()
        check(lib().sld_dev_autotune(self.h, float(dt), int(scheme), float(phi), int(nsteps),
                                     _p(best), C.byref(ms)), "sld_dev_autotune")
        return {"elems_per_block": int(best[0]), "xcd_remap": int(best[1]),
                "persistent": int(best[2]), "assembly": ("scatter", "gather")[int(best[3])],
                "ms_per_step": ms.value}

    def step(self, dt, nsteps=1, scheme=0, phi=1.033):
        check(lib().sld_dev_step(self.h, float(dt), int(scheme), float(phi), int(nsteps)),
              "sld_dev_step")

    def prepare(self, dt, nsteps=1, scheme=0, phi=1.033):
        """Build the captured graphs step(dt, nsteps, ...) would replay from
        the current state; runs no step."""
        check(lib().sld_dev_prepare(self.h, float(dt), int(scheme), float(phi), int(nsteps)),
              "sld_dev_prepare")

    def sync(self):
        check(lib().sld_dev_sync(self.h), "sld_dev_sync")

    def stream(self):
        return lib().sld_dev_stream(self.h)

    def time_kernels(self, dt, nsteps=4, scheme=0, phi=1.033):
        """HIP-event timing of the predictor, element phase and corrector
        over nsteps eager steps (in surrogate mode also its prepass and
        inference)."""
        out = np.zeros(8)
        check(lib().sld_dev_time_kernels(self.h, float(dt), int(scheme), float(phi), int(nsteps),
                                         _p(out)), "sld_dev_time_kernels")
        keys = ["predict_ms", "element_ms", "correct_ms", "element_launches_per_step",
                "kernel_ms_per_step", "stream_ms_per_step", "sur_prepass_ms", "sur_infer_ms"]
        return dict(zip(keys, out.tolist()))

    def critical_dt(self):
        return lib().sld_dev_critical_dt(self.h)

    def coloring(self):
        nc = C.c_int32()
        check(lib().sld_dev_coloring(self.h, C.byref(nc), None, 0, None), "sld_dev_coloring")
        ptr = np.zeros(nc.value + 1, dtype=np.int32)
        order = np.zeros(self.nelem, dtype=np.int32)
        check(lib().sld_dev_coloring(self.h, C.byref(nc), _p(ptr), nc.value, _p(order)),
              "sld_dev_coloring")
        return ptr, order

    # ----------------------------------------------------------------- fields
    def get(self, field):
        f = {"u": 0, "v": 1, "a": 2, "fint": 3, "mass": 4, "sdv": 5, "stress": 6,
             "flags": 7, "time": 8, "counters": 9, "pk1": 10, "reactions": 11,
             "health": 12}[field]
        n, e = self.npoin, self.nelem
        shape, dt = {0: ((n, 3), np.float64), 1: ((n, 3), np.float64), 2: ((n, 3), np.float64),
                     3: ((n, 3), np.float64), 4: ((n,), np.float64), 5: ((e, 8, 3), np.float64),
                     6: ((e, 8, 9), np.float64), 7: ((e, 8), np.uint32), 8: ((1,), np.float64),
                     9: ((8,), np.uint64), 10: ((e, 8, 9), np.float64),
                     11: ((n, 3), np.float64), 12: ((4,), np.float64)}[f]
        out = np.zeros(shape, dtype=dt)
        check(lib().sld_dev_get(self.h, f, _p(out), out.size), f"sld_dev_get({field})")
        return out

    def set(self, field, value):
        f = {"u": 0, "v": 1, "a": 2, "sdv": 5, "time": 8}[field]
        a = np.ascontiguousarray(value, dtype=np.float64)
        check(lib().sld_dev_set(self.h, f, _p(a), a.size), f"sld_dev_set({field})")

    def health(self):
        """(halt bits, element of the most negative det F or -1, that det F,
        lowest node with a non-finite value or -1); bits 1 = det F <= 0
        (sld_detjac), 2 = non-finite displacement or residual (sld_cvgunk)."""
        h = self.get("health")
        return int(h[0]), int(h[1]), float(h[2]), int(h[3])

    def counters(self):
        c = self.get("counters")
        names = ["kf_AS", "kf_SA", "newton_fail", "eigen_fail", "detF_nonpos",
                 "surrogate_used", "surrogate_fallback"]
        return {k: int(v) for k, v in zip(names, c)}
