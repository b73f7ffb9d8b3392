"""TEST INFRASTRUCTURE ONLY.  Speed calibration of the CPU oracle's SMA law:
the reference's own compiled law (oracle/_ref/ref_sm160 in its "quiet"
mode, amdflang -O2, built by ``make -C oracle ref`` from the reference
sources) against the C restatement (orc_sm160_eval, gcc -O3
-ffp-contract=off) on the same seeded points, both on one thread.  Only
where the reference sources were present to build oracle/_ref.

    python oracle/calibrate_law_speed.py [n_points]

Half of the points are elastic (strains ~2e-3), half transform (strains
~3e-2, from the austenite state); material = the bar's Nitinol row
(E = 60 GPa, nu = 0.33, eL = 0.05, AS 400/500, SA 250/150 MPa).
"""
import os
import subprocess
import sys
import tempfile
import time

os.environ["OMP_NUM_THREADS"] = "1"
import numpy as np  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
from oracle import oracle_py as O  # noqa: E402

PROPS = np.array([60000, 0.33, 0.05, 400, 500, 0, 0, 0, 0, 250, 150, 0, 0, 0, 0, 1, 1, 0],
                 dtype=np.float64)


def points(n, seed=7):
    rng = np.random.default_rng(seed)
    amp = np.where(np.arange(n) < n // 2, 2e-3, 3e-2)
    F = np.eye(3)[None] + amp[:, None, None] * rng.normal(size=(n, 3, 3))
    Fsoa = np.ascontiguousarray(F.transpose(0, 2, 1).reshape(n, 9).T)   # [9, n] col-major
    return Fsoa, np.zeros((3, n))


def main():
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 400_000
    if not O.ref_available():
        print("oracle/_ref/ref_sm160 not built: nothing to calibrate")
        return
    F, sdv = points(n)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            np.int64(n).tofile(f)
            np.ascontiguousarray(np.tile(PROPS, (n, 1))).tofile(f)
            np.ascontiguousarray(F.T).tofile(f)
            np.ascontiguousarray(sdv.T).tofile(f)
        out = subprocess.run([os.path.join(O.REF_DIR, "ref_sm160"), fin, fout, "quiet"],
                             check=True, capture_output=True, text=True).stdout
        t_ref = float([l for l in out.splitlines() if l.startswith("#TIME")][0].split()[1])
        raw = np.fromfile(fout, dtype=np.float64)
    S_ref = raw[3 * n:].reshape(n, 9).T
    O.sm160_eval(PROPS[None], F[:, :1000], sdv[:, :1000])          # load and warm
    t0 = time.perf_counter()
    _, S, fl = O.sm160_eval(PROPS[None], F, sdv)
    t_orc = time.perf_counter() - t0
    diff = np.max(np.abs(S - S_ref)) / np.max(np.abs(S_ref))
    print(f"{n} points ({int(np.count_nonzero(fl & 3))} transforming), one thread: "
          f"reference law {t_ref / n * 1e9:.0f} ns/point, C restatement {t_orc / n * 1e9:.0f} ns/point, "
          f"ratio reference/restatement {t_ref / t_orc:.2f}; max |S - S_ref| / max |S_ref| = {diff:.1e}")


if __name__ == "__main__":
    main()
