"""Run stops on the GPU against the oracle: det F <= 0 at a Gauss point
(sld_elmpre.f90:229-240 -> sld_detjac.f90:93) and NaN or infinite
displacements/residuals (sld_cvgunk.f90:195-199) halt the device run, the
next synchronising call returns SMA_ERR_STATE naming the element or node,
the state is the one the oracle stops with, later steps are refused, and
setting the state again clears the halt.  The point-wise law is compared
with the oracle on inverted, singular and NaN deformation gradients."""
import numpy as np
import pytest

from alya_sma_amd import problems
from alya_sma_amd._lib import HipError
from alya_sma_amd.solid import Solidz, eval_sld_stress_model_160
from conftest import soa
from oracle import oracle_py as O

pytestmark = pytest.mark.gpu


def _bar_with_free_corner():
    p = problems.bar_problem(nx=4, ny=2, nz=2, lx=4.0, ly=2.0, lz=2.0, nsteps=100)
    corner = int(np.argmin(np.linalg.norm(p["coord"], axis=1)))
    p["fixmask"] = p["fixmask"].copy()
    p["fixmask"][corner] = 0
    u0 = np.zeros_like(p["coord"])
    u0[corner] = [1.6, 0.3, 0.2]     # through element 0's far x face
    return p, u0


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("mode,assembly,scheme,nsteps",
                         [(0, "gather", 0, 1), (0, "scatter", 1, 1), (1, "gather", 0, 64),
                          (0, "gather", 1, 16)])
def test_negative_jacobian_halts_like_the_oracle(mode, assembly, scheme, nsteps):
    print(f"negative Jacobian: mode {mode} {assembly} scheme {scheme} nsteps {nsteps}", flush=True)
    p, u0 = _bar_with_free_corner()
    dt = p["run"]["dt"]
    with Solidz(p) as dev:
        dev.set_mode(mode)
        dev.set_assembly(assembly)
        ptr, order = dev.coloring()
        orc = O.StepOracle(p, elem_order=order, color_ptr=ptr)
        orc.u[:] = u0
        dev.set("u", u0)
        assert orc.step(dt, scheme=scheme, n=nsteps) == 1
        dev.step(dt, nsteps=nsteps, scheme=scheme)
        with pytest.raises(HipError, match="negative Jacobian.*element 0"):
            dev.sync()
        code, elem, detf, node = dev.health()
        ocode, oelem, odetf, onode = orc.halt
        assert (code, elem, node) == (ocode, oelem, onode) == (1, 0, -1)
        assert abs(detf - odetf) <= 1e-6 * abs(odetf) + (1e-6 if mode else 0.0)
        # the device stopped where the oracle did: after the element loop of
        # the first step (the predictor's u, the half-step v, a untouched)
        assert _rel(dev.get("u"), orc.u) < 1e-13
        assert np.array_equal(dev.get("a"), orc.a)
        assert dev.get("time")[0] == 0.0
        # a halted run refuses to step; a new state clears the halt
        with pytest.raises(HipError, match="halted"):
            dev.step(dt, nsteps=2, scheme=scheme)
        dev.set("u", np.zeros_like(u0))
        dev.set("v", np.zeros_like(u0))
        dev.step(dt, nsteps=2, scheme=scheme)
        dev.sync()
        assert dev.health()[0] == 0


@pytest.mark.parametrize("nsteps", [1, 64])
def test_non_finite_state_halts_like_the_oracle(nsteps):
    print(f"non-finite state: nsteps {nsteps}", flush=True)
    p = problems.bar_problem(nx=4, ny=2, nz=2, lx=4.0, ly=2.0, lz=2.0, nsteps=100)
    dt = p["run"]["dt"]
    free = np.where(p["fixmask"] == 0)[0]
    v0 = np.zeros_like(p["coord"])
    v0[free[3], 1] = np.nan
    v0[free[1], 2] = np.inf
    with Solidz(p) as dev:
        ptr, order = dev.coloring()
        orc = O.StepOracle(p, elem_order=order, color_ptr=ptr)
        orc.v[:] = v0
        dev.set("v", v0)
        assert orc.step(dt, n=nsteps) == 2
        dev.step(dt, nsteps=nsteps)
        with pytest.raises(HipError, match=f"NaN or infinite .* node {min(free[3], free[1])}"):
            dev.sync()
        assert dev.health() == (2, -1, 0.0, orc.halt[3])
        assert dev.get("time")[0] == 0.0


def test_law_on_inverted_singular_and_nan_gradients():
    props = problems.NITINOL
    Fs = [np.diag([-1.02, 0.99, 1.0]),                       # reflection: det < 0
          np.array([[1.0, 0.5, 0.0], [2.0, 1.0, 0.0], [0.0, 0.0, 1.0]]),   # singular
          np.array([[1.01, np.nan, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]]),
          -np.eye(3) @ np.diag([1.03, 1.0, 0.98])]            # inversion: det < 0
    F = soa(np.array(Fs))
    so = np.zeros((3, F.shape[1]))
    sn_g, S_g, fl_g = eval_sld_stress_model_160(props, F, so)
    sn_o, S_o, fl_o = O.sm160_eval(props[None], F, so)
    np.testing.assert_array_equal(fl_g & 0xff1f, fl_o & 0xff1f)
    assert np.all(fl_g & O.FLAG_DETF_NONPOS)
    # the law itself does not look at det F (b = F F^T is positive
    # semi-definite): the inverted gradients give the oracle's stress
    ok = np.isfinite(S_o).all(axis=0)
    assert ok[0] and ok[3]
    assert np.isfinite(S_g[:, ok]).all()
    assert _rel(S_g[:, ok], S_o[:, ok]) < 1e-10
    assert _rel(sn_g[:, ok], sn_o[:, ok]) < 1e-12 or np.max(np.abs(sn_o[:, ok])) == 0.0
    # a NaN gradient gives no finite stress on either side
    assert not np.isfinite(S_o[:, 2]).all() and not np.isfinite(S_g[:, 2]).all()
