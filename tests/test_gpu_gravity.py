"""Gravity on the GPU (sld_dev_set_gravity: m g at every dof, the element
load of sld_elmmat.f90:152-157) against the CPU oracle: a clamped Nitinol
beam under its own (exaggerated) weight, both schemes; reactions carry the
weight of the prescribed nodes; the energy balance sees gravity's work;
gravity can be removed."""
import numpy as np
import pytest

from alya_sma_amd import problems
from alya_sma_amd.solid import Solidz
from oracle import oracle_py as O

pytestmark = pytest.mark.gpu
TOL = 1e-10
G = np.array([0.0, 3.0e10, -8.0e10])     # mm/s^2, large enough to strain the beam ~1e-4


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def _beam():
    prob = problems.cantilever_problem(nx=24, ny=3, nz=3, lx=6.0, ly=0.75, lz=0.75,
                                       traction=0.0, nsteps=600, E_noise=0.02)
    prob["fext0"], prob["fext_func"] = None, -1
    prob["gravity"] = G.copy()
    return prob


@pytest.mark.parametrize("scheme", [0, 1])
def test_gravity_matches_oracle(scheme):
    prob = _beam()
    r = prob["run"]
    with Solidz(prob) as dev:
        ptr, order = dev.coloring()
        orc = O.StepOracle(prob, elem_order=order, color_ptr=ptr)
        for chunk in range(8):
            if chunk:
                dev.set("u", orc.u)
                dev.set("v", orc.v)
                dev.set("a", orc.a)
                dev.set("sdv", orc.sdv[1])
                dev.set("time", np.array([orc.t]))
            dev.step(r["dt"], nsteps=25, scheme=scheme, phi=r["phi"])
            orc.step(r["dt"], scheme=scheme, phi=r["phi"], n=25)
            if chunk < 2:      # strains ~1e-8 at first: below the digits F keeps
                continue
            a_scale = max(np.max(np.abs(orc.fint)) / np.min(orc.mass), np.max(np.abs(orc.a)))
            errs = dict(u=_rel(dev.get("u"), orc.u), v=_rel(dev.get("v"), orc.v),
                        a=np.max(np.abs(dev.get("a") - orc.a)) / a_scale,
                        fint=_rel(dev.get("fint"), orc.fint))
            print(f"gravity scheme {scheme} step {25 * (chunk + 1)}",
                  {k: f"{x:.1e}" for k, x in errs.items()}, flush=True)
            for k, x in errs.items():
                assert x < TOL, (k, x)
        # reactions: frxid = m g - f_int on prescribed dofs
        fixed = np.stack([(prob["fixmask"] >> c) & 1 for c in range(3)], axis=1).astype(bool)
        want = np.where(fixed, orc.mass[:, None] * G - orc.fint, 0.0)
        got = dev.get("reactions")
        assert np.max(np.abs(got - want)) <= 1e-10 * np.max(np.abs(want))
        assert np.abs(orc.u).max() > 1e-4


def test_gravity_work_in_the_energy_balance():
    prob = _beam()
    r = prob["run"]
    with Solidz(prob) as dev:
        ptr, order = dev.coloring()
        orc = O.StepOracle(prob, elem_order=order, color_ptr=ptr, energy=True)
        dev.set_energy(True)
        dev.step(r["dt"], nsteps=150, scheme=0, phi=r["phi"])
        orc.step(r["dt"], scheme=0, phi=r["phi"], n=150)
        e = dev.energy()
        ref = dict(zip(("allie", "allwk", "allke", "etota"), orc.energy()))
        scale = max(abs(x) for x in ref.values())
        print({k: f"{ref[k]:.6e}/{abs(e[k] - ref[k]) / scale:.1e}" for k in ref}, flush=True)
        assert ref["allwk"] > 0.0
        for k in ref:
            assert abs(e[k] - ref[k]) < TOL * scale, k


def test_gravity_removed_before_stepping_changes_nothing():
    prob = _beam()
    r = prob["run"]
    free = dict(prob)
    free.pop("gravity")
    with Solidz(prob) as a, Solidz(free) as b:
        a.step(r["dt"], nsteps=34, scheme=0, phi=r["phi"])     # captured with gravity
        a.set_gravity(None)
        b.step(r["dt"], nsteps=34, scheme=0, phi=r["phi"])
        u_a = a.get("u")
        a.set("u", b.get("u"))
        a.set("v", b.get("v"))
        a.set("a", b.get("a"))
        a.set("sdv", b.get("sdv"))
        a.set("time", b.get("time"))
        a.step(r["dt"], nsteps=34, scheme=0, phi=r["phi"])
        b.step(r["dt"], nsteps=34, scheme=0, phi=r["phi"])
        np.testing.assert_array_equal(a.get("u"), b.get("u"))
        assert np.abs(u_a).max() > 0.0
