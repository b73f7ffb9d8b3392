"""Point-wise SMA law on the GPU (sma160_eval / sma160_eval_device) against
the CPU oracle and the reference's golden vectors."""
import numpy as np
import pytest

from alya_sma_amd.solid import eval_sld_stress_model_160, eval_sld_stress_model_160_device
from conftest import soa
from oracle import oracle_py as O

pytestmark = pytest.mark.gpu
TOL = 1e-10     # FP64 relative tolerance on stress


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def test_law_fp64_golden(golden_sm160):
    g = golden_sm160
    print("law fp64 golden n =", g["F"].shape[1], flush=True)
    sn, S, fl = eval_sld_stress_model_160(g["mats"], g["F"], g["sdv_old"], mat_of_gp=g["mat"])
    ok = np.isfinite(g["S"]).all(axis=0)
    scale = np.maximum(np.max(np.abs(g["S"][:, ok]), axis=0), 1e-300)
    per_pt = np.max(np.abs(S[:, ok] - g["S"][:, ok]), axis=0) / scale
    assert per_pt.max() < TOL
    zero = np.max(np.abs(g["S"][:, ok]), axis=0) == 0          # F = I: zero stress
    assert np.max(np.abs(S[:, ok][:, zero]), initial=0.0) < 1e-12
    np.testing.assert_allclose(sn[0, ok], g["sdv_new"][0, ok], rtol=0, atol=1e-12)
    assert _rel(sn[1:, ok], g["sdv_new"][1:, ok]) < 1e-13
    _, _, flo = O.sm160_eval(g["mats"], g["F"], g["sdv_old"], mat_of_gp=g["mat"])
    np.testing.assert_array_equal(fl & 0xff1f, flo & 0xff1f)               # branch, fail, iters
    assert np.all(((g["warn"] & 2) > 0) == ((fl & 4) > 0))
    assert np.all(((g["warn"] & 4) > 0) == ((fl & 8) > 0))


def test_law_fp32_error_on_realistic_points():
    rng = np.random.default_rng(7)
    n = 20000
    props = np.array([60000, 0.33, 0.05, 400, 500, 0, 0, 0, 0, 250, 150, 0, 0, 0, 0, 1, 1, 0.0])
    # paths from rest: the history comes from the exact law at the previous state
    F0 = np.eye(3)[None] + rng.normal(0, 0.02, (n, 3, 3))
    F1 = F0 + rng.normal(0, 0.001, (n, 3, 3))
    so, _, _ = O.sm160_eval(props[None], soa(F0), np.zeros((3, n)))
    sn_o, S_o, fl_o = O.sm160_eval(props[None], soa(F1), so)
    sn_g, S_g, fl_g = eval_sld_stress_model_160(props, soa(F1), so, precision=1)
    err = np.linalg.norm(S_g - S_o) / np.linalg.norm(S_o)
    flip = np.mean((fl_g & 3) != (fl_o & 3))
    print(f"fp32 law: rel L2 stress err {err:.2e}, branch flips {flip:.2e}", flush=True)
    assert err < 1e-4 and flip < 1e-3


def test_law_device_pointers(golden_sm160):
    from alya_sma_amd.device import DeviceArray, synchronize
    g = golden_sm160
    ok = np.isfinite(g["S"]).all(axis=0)
    F = DeviceArray.from_host(g["F"][:, ok])
    so = DeviceArray.from_host(g["sdv_old"][:, ok])
    props = DeviceArray.from_host(g["mats"])
    mat = DeviceArray.from_host(g["mat"][ok].astype(np.int32))
    n = F.shape[1]
    sn = DeviceArray((3, n), np.float64)
    S = DeviceArray((9, n), np.float64)
    eval_sld_stress_model_160_device(n, props.shape[0], props.ptr, F.ptr, so.ptr, sn.ptr, S.ptr,
                                     mat_ptr=mat.ptr)
    synchronize()
    assert _rel(S.to_host(), g["S"][:, ok]) < TOL


def test_law_empty_and_bad_args():
    sn, S, fl = eval_sld_stress_model_160(np.zeros(18), np.zeros((9, 0)), np.zeros((3, 0)))
    assert S.shape == (9, 0)
    from alya_sma_amd._lib import HipError
    with pytest.raises(HipError, match="material index"):
        eval_sld_stress_model_160(np.zeros(18), np.eye(3).reshape(9, 1), np.zeros((3, 1)),
                                  mat_of_gp=np.array([3]))
