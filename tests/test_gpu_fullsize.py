"""Parity at the benchmark sizes: the cantilever 400x20x20 (1.28 M Gauss
points), the 256-specimen batch (21.7 M) and the SMA tube in tissue (48.4 M).
The device runs into the loading, its state is copied into the FP64 CPU
oracle (run in the device's colour order, OpenMP over each colour), and both
take the next steps: every nodal field, the history and the branch flags
agree to 1e-10 in FP64, and the FP32 mode's stress is within 1e-4 (relative
L2) of the FP64 step from the same state; the cantilever also with the
Tchamwa-Wielgosz scheme."""
import numpy as np
import pytest

from alya_sma_amd import problems
from alya_sma_amd.solid import Solidz
from oracle import oracle_py as O

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def _load_state(orc, dev):
    orc.u[:] = dev.get("u")
    orc.v[:] = dev.get("v")
    orc.a[:] = dev.get("a")
    sdv = dev.get("sdv")
    orc.sdv[0] = sdv
    orc.sdv[1] = sdv
    orc.s.t = float(dev.get("time")[0])


# Device steps before the comparison.  The state handed over must lie where
# the reference's law is defined: its eigen-solver underflows, and the law
# returns NaN, when the off-diagonals of b lie between ~1e-162 and ~1e-77
# (tests/test_oracle_law.py; the oracle restates that bit for bit).  Ahead
# of a loading wave every strain component falls through that band (on the
# 256 specimens the wave needs up to 1956 steps to cross the longest one),
# and the device's FP64 step, whose series strain adds no rounding noise to
# a zero shear, also keeps shear strains of that size in the symmetric
# uniaxial specimens.  The handed-over displacement therefore carries a
# perturbation of 1e-15 of its largest value (a generic state, the same on
# both sides), which lifts every strain component above ~1e-14.
CASES = {
    "cantilever": (lambda: problems.cantilever_problem(), 1500),
    "specimens": (lambda: problems.batched_specimens(), 1000),
    "cylinder": (lambda: problems.cylinder_problem(), 3000),
}


@pytest.mark.parametrize("case,scheme", [(c, 0) for c in CASES] + [("cantilever", 1)])
def test_full_size_step_matches_oracle(case, scheme):
    """scheme 0 = central difference, 1 = Tchamwa-Wielgosz."""
    build, warm = CASES[case]
    prob = build()
    r = prob["run"]
    n_gp = 8 * prob["lnods"].shape[0]
    print(f"{case} scheme {scheme}: {n_gp} Gauss points, {warm} device steps, then 2 steps on both",
          flush=True)
    with Solidz(prob) as dev:
        dev.set_store(True)
        dev.step(r["dt"], nsteps=warm, scheme=scheme, phi=r["phi"])
        u0 = dev.get("u")
        rng = np.random.default_rng(5)
        dev.set("u", u0 + 1e-15 * np.abs(u0).max() * rng.standard_normal(u0.shape))
        ptr, order = dev.coloring()
        orc = O.StepOracle(prob, elem_order=order, color_ptr=ptr)
        _load_state(orc, dev)
        dev.step(r["dt"], nsteps=2, scheme=scheme, phi=r["phi"])
        orc.step(r["dt"], scheme=scheme, phi=r["phi"], n=2)
        u, f = dev.get("u"), dev.get("fint")
        a_scale = max(np.max(np.abs(orc.fint)) / np.min(orc.mass), np.max(np.abs(orc.a)))
        S_scale = max(np.max(np.abs(orc.S)), 1e-4 * orc.E_max)
        errs = dict(u=_rel(u, orc.u), v=_rel(dev.get("v"), orc.v),
                    a=np.max(np.abs(dev.get("a") - orc.a)) / a_scale, fint=_rel(f, orc.fint),
                    S=np.max(np.abs(dev.get("stress") - orc.S)) / S_scale)
        sdv = dev.get("sdv")
        fl = dev.get("flags")
        moving = np.abs(orc.u).max()
        transforming = int(np.count_nonzero(orc.flags & 3))
        print(f"  errors {({k: f'{x:.1e}' for k, x in errs.items()})}, max |u| {moving:.3e}, "
              f"transforming points {transforming}", flush=True)
        if not all(x < TOL for x in errs.values()):
            du = np.abs(u - orc.u).max(axis=1)
            p = int(np.argmax(du))
            print(f"  device health {dev.health()}, oracle halt {orc.halt}, time device "
                  f"{dev.get('time')[0]!r} oracle {orc.t!r}; largest u difference at node {p}: "
                  f"device {u[p]} oracle {orc.u[p]}, fixmask {prob['fixmask'][p]}, "
                  f"non-finite oracle v {int(np.count_nonzero(~np.isfinite(orc.v)))}, "
                  f"S {int(np.count_nonzero(~np.isfinite(orc.S)))}, device v "
                  f"{int(np.count_nonzero(~np.isfinite(dev.get('v'))))}", flush=True)
        assert moving > 0.0
        for k, x in errs.items():
            assert x < TOL, (case, k, x)
        np.testing.assert_array_equal(fl & 3, orc.flags & 3)
        assert _rel(sdv[..., 1:], orc.sdv[0, ..., 1:]) < TOL
        assert np.max(np.abs(sdv[..., 0] - orc.sdv[0, ..., 0])) < 1e-9

        # FP32 mode from the same state: one step against the FP64 step
        S64 = orc.S.copy()
        _load_state(orc, dev)       # the device state after the two steps
        orc.step(r["dt"], scheme=scheme, phi=r["phi"], n=1)
        dev.set_mode(1)
        dev.step(r["dt"], nsteps=1, scheme=scheme, phi=r["phi"])
        S32 = dev.get("stress")
        l2 = np.linalg.norm(S32 - orc.S) / np.linalg.norm(orc.S)
        print(f"  FP32 one-step stress error {l2:.2e} (relative L2)", flush=True)
        assert l2 < 1e-4
        assert np.abs(S64).max() > 0.0


def test_surrogate_mode_on_the_cantilever_against_oracle():
    """Surrogate mode (FP32 kinematics, trained network on single-branch
    transforming points, exact return map elsewhere) on the 400x20x20
    cantilever: the device runs into the transforming phase in FP64, its
    state goes to the oracle and back to the device, and one surrogate step
    stays within 1e-3 (relative L2 stress) of the oracle's FP64 step, with
    accepted points present and their history within the acceptance
    tolerance of the exact map's."""
    from alya_sma_amd import surrogate as S
    net = S.load_default()
    assert net is not None, "the committed network is missing"
    prob = problems.cantilever_problem()
    r = prob["run"]
    print("cantilever surrogate mode: 2000 FP64 steps, then one surrogate step", flush=True)
    with Solidz(prob) as dev:
        dev.set_store(True)
        dev.step(r["dt"], nsteps=2000)
        ptr, order = dev.coloring()
        orc = O.StepOracle(prob, elem_order=order, color_ptr=ptr)
        _load_state(orc, dev)
        dev.set_surrogate(net, 1e-4)
        dev.set_mode(2)
        c0 = dev.counters()
        dev.step(r["dt"], nsteps=1)
        orc.step(r["dt"], n=1)
        c1 = dev.counters()
        Sd = dev.get("stress")
        l2 = np.linalg.norm(Sd - orc.S) / np.linalg.norm(orc.S)
        used = c1["surrogate_used"] - c0["surrogate_used"]
        fall = c1["surrogate_fallback"] - c0["surrogate_fallback"]
        transforming = int(np.count_nonzero(orc.flags & 3))
        fl = dev.get("flags")
        # accepted points on the FP64 step's branch (an FP32 loading function
        # within rounding of a threshold may take the other branch)
        acc = ((fl & 0x20) > 0) & ((fl & 3) == (orc.flags & 3))
        dxs = np.abs(dev.get("sdv")[..., 0] - orc.sdv[0, ..., 0])
        print(f"  stress rel L2 {l2:.2e}, transforming {transforming}, accepted {used}, "
              f"fallback {fall}, max |dxS| accepted {dxs[acc].max() if acc.any() else 0:.2e}",
              flush=True)
        assert transforming > 0 and used > 0
        assert l2 < 1e-3
        assert dxs[acc].max() <= 1e-4 + 1e-6
