"""Whole-run parity of the FP64 step on the cantilever 400x20x20 (5000 steps):
at the start of every window of W steps the oracle takes the device's state,
both run W steps, and the largest differences are printed (u, v, f_int
relative to their maxima, a relative to max|f_int|/min m, S relative to
max(|S|, E 1e-4)) with the number of points whose branch flags differ.
Needs the GPU; about three minutes.

    python tests/run_full_trajectory_parity.py [W=25] [steps=5000]
"""
import os, sys, time
sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import numpy as np
from alya_sma_amd import problems
from alya_sma_amd.solid import Solidz
from oracle import oracle_py as O
prob = problems.cantilever_problem()
r = prob["run"]
W = int(sys.argv[1]) if len(sys.argv) > 1 else 25
N = int(sys.argv[2]) if len(sys.argv) > 2 else 5000
rel = lambda a, b: np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)
worst = {}
t0 = time.time()
with Solidz(prob) as dev:
    dev.set_store(True)
    ptr, order = dev.coloring()
    orc = O.StepOracle(prob, elem_order=order, color_ptr=ptr)
    done = 0
    while done < N:
        # the oracle takes the device's state, both run W steps
        orc.u[:] = dev.get("u"); orc.v[:] = dev.get("v"); orc.a[:] = dev.get("a")
        sdv = dev.get("sdv"); orc.sdv[0] = sdv; orc.sdv[1] = sdv
        orc.s.t = float(dev.get("time")[0])
        dev.step(r["dt"], nsteps=W, scheme=0, phi=r["phi"])
        orc.step(r["dt"], scheme=0, phi=r["phi"], n=W)
        done += W
        a_scale = max(np.max(np.abs(orc.fint)) / np.min(orc.mass), np.max(np.abs(orc.a)))
        S_scale = max(np.max(np.abs(orc.S)), 1e-4 * orc.E_max)
        e = dict(u=rel(dev.get("u"), orc.u), v=rel(dev.get("v"), orc.v),
                 a=np.max(np.abs(dev.get("a") - orc.a)) / a_scale, fint=rel(dev.get("fint"), orc.fint),
                 S=np.max(np.abs(dev.get("stress") - orc.S)) / S_scale)
        fl = int(np.count_nonzero((dev.get("flags") & 3) != (orc.flags & 3)))
        tr = int(np.count_nonzero(orc.flags & 3))
        for k, x in e.items():
            worst[k] = max(worst.get(k, 0.0), x)
        worst["flag_mismatch"] = worst.get("flag_mismatch", 0) + fl
        print(f"step {done}: " + " ".join(f"{k} {x:.1e}" for k, x in e.items())
              + f" flags differ {fl} transforming {tr} ({time.time() - t0:.0f} s)", flush=True)
print("worst", worst, flush=True)
