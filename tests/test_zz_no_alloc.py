"""sma160_eval_noalloc (staging through the library's fixed device array, no
allocation) against sma160_eval (staging through buffers it allocates per
call) and the CPU oracle: two calls in a row, the second smaller, in FP64 and
FP32 kinematics; sizes above the maximum are refused."""
import ctypes as C

import numpy as np
import pytest

from alya_sma_amd import _lib
from alya_sma_amd.problems import NITINOL
from conftest import soa
from oracle import oracle_py as O

pytestmark = pytest.mark.gpu

MAX_GP = 16384   # SMA160_NOALLOC_MAX_GP of include/alya_sma.h


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _bind():
    L = _lib.lib()
    p, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
    L.sma160_eval_noalloc.argtypes = [i64, i32, p, p, p, p, p, p, p, p, p, i32]
    L.sma160_eval_noalloc.restype = C.c_int
    return L


def _inputs(n, seed, with_mat):
    """Points on a load path from rest (about half transforming), two
    materials (NITINOL and a 10% stiffer variant), per-point moduli."""
    rng = np.random.default_rng(seed)
    mats = np.stack([NITINOL, NITINOL * np.r_[1.1, 1.0, np.ones(16)]])
    F0 = np.eye(3)[None] + rng.normal(0, 0.012, (n, 3, 3))
    F1 = F0 + rng.normal(0, 0.001, (n, 3, 3))
    mat = rng.integers(0, 2, n).astype(np.int32) if with_mat else None
    E = (60000.0 * (1.0 + rng.uniform(-0.02, 0.02, n))) if with_mat else None
    so, _, _ = O.sm160_eval(mats, soa(F0), np.zeros((3, n)), mat_of_gp=mat, E_of_gp=E)
    return mats, np.ascontiguousarray(soa(F1)), np.ascontiguousarray(so), mat, E


def _call(fn, mats, F, so, mat, E, precision):
    n = F.shape[1]
    sn = np.zeros((3, n))
    S = np.zeros((9, n))
    fl = np.zeros(n, dtype=np.uint32)
    rc = fn(n, mats.shape[0], _p(mats), _p(mat), _p(E), _p(F), None, _p(so), _p(sn), _p(S),
            _p(fl), precision)
    assert rc == 0, _lib.lib().sma_last_error().decode()
    return sn, S, fl


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_matches_allocating_entry_point(precision):
    L = _bind()
    # first call near the maximum with materials and moduli per point, the
    # second smaller and without them (material 0, props' modulus)
    for n, seed, with_mat in ((MAX_GP - 5, 11, True), (3001, 12, False)):
        print(f"noalloc precision {precision} n {n} mat {with_mat}", flush=True)
        mats, F, so, mat, E = _inputs(n, seed, with_mat)
        sn_a, S_a, fl_a = _call(L.sma160_eval, mats, F, so, mat, E, precision)
        sn_b, S_b, fl_b = _call(L.sma160_eval_noalloc, mats, F, so, mat, E, precision)
        # the same kernel on the same inputs: bitwise equal
        assert np.array_equal(S_a, S_b) and np.array_equal(sn_a, sn_b)
        assert np.array_equal(fl_a, fl_b)
        ntr = int(np.count_nonzero(fl_b & 3))
        assert ntr > n // 10, ntr   # the transforming branches are exercised
        if precision == 0:
            sn_o, S_o, fl_o = O.sm160_eval(mats, F, so, mat_of_gp=mat, E_of_gp=E)
            scale = np.maximum(np.max(np.abs(S_o), axis=0), 1e-300)
            assert np.max(np.max(np.abs(S_b - S_o), axis=0) / scale) < 1e-10
            assert np.array_equal(fl_b & 3, fl_o & 3)


def test_noalloc_refuses_sizes_above_the_maximum():
    L = _bind()
    F = np.zeros((9, 1))
    z = np.zeros((3, 1))
    S = np.zeros((9, 1))
    rc = L.sma160_eval_noalloc(MAX_GP + 1, 1, _p(NITINOL), None, None, _p(F), None, _p(z), _p(z),
                               _p(S), None, 0)
    assert rc == -1 and b"exceeds the maximum" in L.sma_last_error()
    props = np.tile(NITINOL, (65, 1))
    rc = L.sma160_eval_noalloc(1, 65, _p(props), None, None, _p(F), None, _p(z), _p(z), _p(S),
                               None, 0)
    assert rc == -1 and b"exceeds the maximum" in L.sma_last_error()
