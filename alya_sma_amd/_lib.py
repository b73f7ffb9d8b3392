"""ctypes binding of the C ABI in include/alya_sma.h (libalya_sma_amd.so).

The library is built in-tree by ``alya_sma_amd.build.build()``.  Loading
fails loudly when it is missing: there is no CPU fallback.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libalya_sma_amd.so")

SMA_FLAG_KF_AS = 0x01
SMA_FLAG_KF_SA = 0x02
SMA_FLAG_NEWTON_FAIL = 0x04
SMA_FLAG_EIGEN_FAIL = 0x08
SMA_FLAG_DETF_NONPOS = 0x10
SMA_FLAG_SURR_USED = 0x20
SMA_FLAG_SURR_FALLBACK = 0x40
SMA_FLAG_ITER_SHIFT = 8

SLD_MODE_FP64 = 0
SLD_MODE_FP32 = 1
SLD_MODE_SURROGATE = 2

SLD_FIELD_U, SLD_FIELD_V, SLD_FIELD_A, SLD_FIELD_FINT, SLD_FIELD_MASS = 0, 1, 2, 3, 4
SLD_FIELD_SDV, SLD_FIELD_STRESS, SLD_FIELD_FLAGS, SLD_FIELD_TIME, SLD_FIELD_COUNTERS = 5, 6, 7, 8, 9

class SldMeshDesc(C.Structure):
    _fields_ = [("npoin", C.c_int64), ("nelem", C.c_int64),
                ("lnods", C.c_void_p), ("coord", C.c_void_p)]


class SldMaterialDesc(C.Structure):
    _fields_ = [("nmat", C.c_int32), ("props", C.c_void_p), ("rho", C.c_void_p),
                ("mat_of_elem", C.c_void_p), ("E_of_elem", C.c_void_p), ("law", C.c_void_p)]


class SldFaceDesc(C.Structure):
    _fields_ = [("nface", C.c_int64), ("lnodb", C.c_void_p), ("code", C.c_void_p),
                ("value", C.c_void_p), ("func", C.c_void_p), ("transient", C.c_int32),
                ("follower", C.c_int32)]


class SldBcDesc(C.Structure):
    _fields_ = [("fixmask", C.c_void_p), ("func_of_node", C.c_void_p),
                ("nfunc", C.c_int32), ("func_maxpts", C.c_int32),
                ("func_npts", C.c_void_p), ("func_smooth", C.c_void_p),
                ("func_table", C.c_void_p), ("fext0", C.c_void_p),
                ("fext_func", C.c_int32)]


class HipError(RuntimeError):
    pass


_lib = None


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HipError(f"{LIB_PATH} is missing: build it with "
                       "`python -c 'import __graft_entry__ as g; g.build()'`")
    L = C.CDLL(LIB_PATH)
    p, i32, i64, d = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    L.sma_last_error.restype = C.c_char_p
    L.sma160_eval.argtypes = [i64, i32, p, p, p, p, p, p, p, p, p, i32]
    L.sma160_eval_device.argtypes = [i64, i32, p, p, p, p, p, p, p, p, p, i32, p]
    L.sld_dev_create.argtypes = [p, p, p, C.POINTER(p)]
    L.sld_dev_set_mode.argtypes = [p, i32]
    L.sld_dev_set_store.argtypes = [p, i32]
    L.sld_dev_step.argtypes = [p, d, i32, d, i64]
    L.sld_dev_prepare.argtypes = [p, d, i32, d, i64]
    L.sld_dev_time_kernels.argtypes = [p, d, i32, d, i32, p]
    L.sld_dev_critical_dt.argtypes = [p]
    L.sld_dev_critical_dt.restype = d
    L.sld_dev_get.argtypes = [p, i32, p, i64]
    L.sld_dev_set.argtypes = [p, i32, p, i64]
    L.sld_dev_coloring.argtypes = [p, p, p, i32, p]
    L.sld_dev_set_tile.argtypes = [p, i32, i32, i32]
    L.sld_dev_set_assembly.argtypes = [p, i32]
    L.sld_dev_set_gravity.argtypes = [p, C.POINTER(C.c_double)]
    L.sld_dev_set_energy.argtypes = [p, i32]
    L.sld_dev_set_damping.argtypes = [p, p]
    L.sld_dev_set_body_force.argtypes = [p, i32]
    L.sld_dev_set_faces.argtypes = [p, p]
    L.sld_dev_energy.argtypes = [p, p]
    L.sld_dev_autotune.argtypes = [p, d, i32, d, i32, p, p]
    L.sld_dev_stream.argtypes = [p]
    L.sld_dev_stream.restype = p
    L.sld_dev_sync.argtypes = [p]
    L.sld_dev_destroy.argtypes = [p]
    L.sma_device_malloc.argtypes = [C.POINTER(p), i64]
    L.sma_device_free.argtypes = [p]
    L.sma_memcpy.argtypes = [p, p, i64, i32]
    L.sma_device_synchronize.argtypes = []
    L.sma_device_count.argtypes = [C.POINTER(i32)]
    L.sma_copy_probe.argtypes = [p, p, i64, p]
    L.sld_mesh_box.argtypes = [i32, i32, i32, d, d, d, p, p]
    L.sld_color_elements.argtypes = [i64, i64, p, p]
    L.sld_color_elements.restype = i32
    L.sma_mlp_create.argtypes = [C.c_uint64, C.POINTER(p)]
    L.sma_mlp_destroy.argtypes = [p]
    L.sma_mlp_set_params.argtypes = [p, p, p]
    L.sma_mlp_get_params.argtypes = [p, p, p]
    L.sma_mlp_set_loss_weights.argtypes = [p, p]
    L.sma_mlp_forward_device.argtypes = [p, i64, p, p, p]
    L.sma_mlp_grad_device.argtypes = [p, i64, p, p, p, p, p, p]
    L.sma_mlp_train_step_device.argtypes = [p, i64, p, p, d, p, p]
    L.sma_mlp_gen_samples_device.argtypes = [i64, C.c_uint64, p, d, d, d, p, p, p, p]
    if hasattr(L, "sld_dev_set_surrogate"):
        L.sld_dev_set_surrogate.argtypes = [p, p, d]
    _lib = L
    return L


def check(rc, what=""):
    if rc != 0:
        msg = lib().sma_last_error().decode(errors="replace")
        raise HipError(f"{what}: status {rc}: {msg}")
    return rc
