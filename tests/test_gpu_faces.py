"""Neumann loads on boundary faces on the GPU (sld_dev_set_faces, the
sld_bouope of sld_bouope.f90) against the CPU oracle: follower pressure with
a transient function and a constant dead traction, both schemes, both
assemblies; the energy balance sees the face work; faces can be replaced
and removed between steps."""
import numpy as np
import pytest

from alya_sma_amd import problems
from alya_sma_amd.solid import Solidz
from oracle import oracle_py as O

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def _beam(load, nsteps=600):
    """Nitinol cantilever 24x3x3 over 6 x 0.75 x 0.75 mm clamped at x = 0.
    load "pressure": follower pressure on the top face z = lz, magnitude
    -p(t) (pushing in) from a smooth 0 -> 1 -> 0 function (transient);
    load "traction": constant dead traction (40, 25, 0) MPa on x = lx."""
    prob = problems.cantilever_problem(nx=24, ny=3, nz=3, lx=6.0, ly=0.75, lz=0.75,
                                       traction=0.0, nsteps=nsteps, E_noise=0.02)
    prob["fext0"], prob["fext_func"] = None, -1
    if load == "pressure":
        sel = lambda c: np.abs(c[:, 2] - 0.75) < 1e-9
        fb, _ = problems.boundary_faces(prob["coord"], prob["lnods"], sel)
        tab = prob["func_table"].copy()
        tab[0, :, 0] *= -60.0                  # column x: the pressure value
        prob["func_table"] = tab
        prob["faces"] = dict(lnodb=fb, code=np.full(len(fb), 3, np.int32),
                             value=np.zeros((len(fb), 3)), func=np.zeros(len(fb), np.int32),
                             transient=True, follower=True)
    else:
        sel = lambda c: np.abs(c[:, 0] - 6.0) < 1e-9
        fb, _ = problems.boundary_faces(prob["coord"], prob["lnods"], sel)
        prob["faces"] = dict(lnodb=fb, code=np.full(len(fb), 6, np.int32),
                             value=np.tile([40.0, 25.0, 0.0], (len(fb), 1)), func=None,
                             transient=False, follower=False)
    return prob


def _compare(dev, orc, label):
    a_scale = max(np.max(np.abs(orc.fint)) / np.min(orc.mass), np.max(np.abs(orc.a)))
    errs = dict(u=_rel(dev.get("u"), orc.u), v=_rel(dev.get("v"), orc.v),
                a=np.max(np.abs(dev.get("a") - orc.a)) / a_scale, fint=_rel(dev.get("fint"), orc.fint))
    print(label, {k: f"{x:.1e}" for k, x in errs.items()}, flush=True)
    for k, x in errs.items():
        assert x < TOL, (label, k, x)


@pytest.mark.parametrize("load", ["pressure", "traction"])
@pytest.mark.parametrize("scheme", [0, 1])
@pytest.mark.parametrize("assembly", ["gather", "scatter"])
def test_face_loads_match_oracle(load, scheme, assembly):
    prob = _beam(load)
    r = prob["run"]
    print(f"faces {load} scheme {scheme} {assembly}", flush=True)
    with Solidz(prob) as dev:
        dev.set_assembly(assembly)
        ptr, order = dev.coloring()
        orc = O.StepOracle(prob, elem_order=order, color_ptr=ptr)
        for chunk in range(12):
            if chunk:
                dev.set("u", orc.u)
                dev.set("v", orc.v)
                dev.set("a", orc.a)
                dev.set("sdv", orc.sdv[1])
                dev.set("time", np.array([orc.t]))
            dev.step(r["dt"], nsteps=25, scheme=scheme, phi=r["phi"])
            orc.step(r["dt"], scheme=scheme, phi=r["phi"], n=25)
            # the first steps of the smooth ramp strain the beam by ~1e-8, where
            # the log strain keeps only ~1e-8 of its digits: compare once the
            # load is established
            if chunk >= 3:
                _compare(dev, orc, f"  step {25 * (chunk + 1)}")
        assert np.abs(orc.u).max() > 1e-4


def test_face_work_in_the_energy_balance():
    prob = _beam("pressure")
    r = prob["run"]
    with Solidz(prob) as dev:
        ptr, order = dev.coloring()
        orc = O.StepOracle(prob, elem_order=order, color_ptr=ptr, energy=True)
        dev.set_energy(True)
        dev.step(r["dt"], nsteps=150, scheme=1, phi=r["phi"])
        orc.step(r["dt"], scheme=1, phi=r["phi"], n=150)
        e = dev.energy()
        ref = dict(zip(("allie", "allwk", "allke", "etota"), orc.energy()))
        scale = max(abs(x) for x in ref.values())
        print({k: f"{ref[k]:.6e}/{abs(e[k] - ref[k]) / scale:.1e}" for k in ref}, flush=True)
        assert ref["allwk"] > 0.0
        for k in ref:
            assert abs(e[k] - ref[k]) < TOL * scale, k


def test_faces_replaced_and_removed_between_steps():
    prob = _beam("traction")
    r = prob["run"]
    faces = prob.pop("faces")
    with Solidz(prob) as dev:
        orc_free = O.StepOracle(prob)
        dev.step(r["dt"], nsteps=40, scheme=0, phi=r["phi"])   # no faces: graph captured
        orc_free.step(r["dt"], scheme=0, phi=r["phi"], n=40)
        np.testing.assert_array_equal(dev.get("u"), 0.0 * orc_free.u)
        dev.set_faces(faces)
        dev.step(r["dt"], nsteps=40, scheme=0, phi=r["phi"])
        assert np.abs(dev.get("u")).max() > 0.0
        dev.set_faces(None)
        u1 = dev.get("u")
        dev.step(r["dt"], nsteps=2, scheme=0, phi=r["phi"])
        assert np.abs(dev.get("fint")).max() > 0.0                # still moving, no load
        bad = dict(faces, code=np.full(len(faces["code"]), 4, np.int32))
        with pytest.raises(RuntimeError):
            dev.set_faces(bad)
        assert np.isfinite(u1).all()
