"""Writes the golden vectors under tests/golden/ from the reference's own
compiled routines (oracle/_ref, built by ``make -C oracle ref`` from the
reference sources; only possible in a container that holds them).

    python tests/golden/gen_golden.py

sm160_golden.npz  -- the SMA law (sld_stress_model_160.f90, flagt=0) on
    seeded load/unload/reload paths (rotated general stretches, exact
    uniaxial stretches with repeated eigenvalues, pure rotations, identity),
    on random histories that reach both-flag and clamp cases, on points that
    make the Newton loop hit its 65-iteration cap, and on NaN inputs that make
    DSYEVQ3 give up.  Inputs: material table, per-point material index, F
    (SoA 9 x n, column-major per point), sdv_old (3 x n).  Outputs: sdv_new,
    S, and the warnings the reference printed per point.
hex8_golden.npz   -- shape3/gauss_qua tables and jacobi/invmtx kinematics on
    seeded distorted HEX08 elements with random displacements.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle_py as O  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))

# Material table: [E, nu, eL, sS_AS, sF_AS, C_AS, Tf_AS, Ts_AS, sN_AS(unused),
#                  sS_SA, sF_SA, C_SA, Tf_SA, Ts_SA, alpha, B_AS, B_SA, unused]
MATS = np.array([
    # config materials: E=60 GPa, nu=0.33, eL=0.05, AS 400/500, SA 250/150 MPa
    [60000, 0.33, 0.05, 400, 500, 0, 0, 0, 0, 250, 150, 0, 0, 0, 0, 1, 1, 0],
    # the NITINOL line of tests/uniaxial_test/sm154_d1C.sld.dat:21-22
    [1000, 0.3, 0.1, 90, 150, 0, 0, 0, 700, 70, 30, 0, 0, 0, 0, 10, 10, 0],
    # volumetric coupling (alpha != 0) and Clausius-Clapeyron slopes (inert T)
    [55000, 0.30, 0.045, 380, 480, 6.5, 280, 300, 0, 240, 140, 6.5, 250, 270, 0.02, 1, 1, 0],
    # jittered config material (+-10%)
    [63000, 0.31, 0.052, 420, 530, 0, 0, 0, 0, 235, 160, 0, 0, 0, 0, 1, 1, 0],
    # reverse threshold above the forward one: lets kf_AS and kf_SA hold
    # together (the 'WARNING 1' branch, sld_stress_model_160.f90:377-379)
    [50000, 0.30, 0.05, 200, 300, 0, 0, 0, 0, 350, 250, 0, 0, 0, 0, 1, 1, 0],
], dtype=np.float64)


def rot(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a*a+b*b-c*c-d*d, 2*(b*c-a*d), 2*(b*d+a*c)],
                     [2*(b*c+a*d), a*a-b*b+c*c-d*d, 2*(c*d-a*b)],
                     [2*(b*d-a*c), 2*(c*d+a*b), a*a-b*b-c*c+d*d]])


def soa(Fs):
    """[n,3,3] (row, col) -> SoA [9,n] column-major per point."""
    return np.ascontiguousarray(Fs.transpose(0, 2, 1).reshape(-1, 9).T)


def run_paths(rng):
    """Load/unload/reload paths; each step's history is the previous output."""
    npath, nstep = 96, 24
    prof = np.concatenate([np.linspace(0, 1, 10), np.linspace(1, 0.1, 8)[1:], np.linspace(0.1, 0.8, 8)[1:]])[:nstep]
    Fpath = np.zeros((npath, nstep, 3, 3))
    mat = np.zeros(npath, dtype=np.int32)
    for p in range(npath):
        kind = p % 4
        mat[p] = (p // 4) % len(MATS)
        amp = rng.uniform(0.02, 0.09) * (1.0 if mat[p] != 1 else 2.0)
        R = rot(rng)
        if kind == 0:      # general rotated stretch with shear
            H = rng.normal(size=(3, 3))
            H = 0.5 * (H + H.T) + 0.2 * (H - H.T)
            H *= amp / np.linalg.norm(H)
        elif kind == 1:    # exact uniaxial stretch: two equal eigenvalues
            nu = MATS[mat[p], 1]
            H = np.diag([1.0, -nu, -nu]) * amp
            R = np.eye(3)
        elif kind == 2:    # rotated uniaxial (repeated eigenvalues, rotated basis)
            nu = MATS[mat[p], 1]
            H = np.diag([1.0, -nu, -nu]) * amp
        else:              # simple shear
            H = np.zeros((3, 3)); H[0, 1] = 1.5 * amp
        for k in range(nstep):
            Fpath[p, k] = R @ (np.eye(3) + prof[k] * H)
    props_pp = MATS[mat]
    sdv = np.zeros((3, npath))
    rows = []
    for k in range(nstep):
        Fk = soa(Fpath[:, k])
        sn, S, w = O.ref_sm160(props_pp, Fk, sdv)
        rows.append((mat.copy(), Fk, sdv.copy(), sn, S, w))
        sdv = sn
    return rows


def run_random(rng, n=1200):
    mat = rng.integers(0, len(MATS), n).astype(np.int32)
    Fs = np.eye(3) + rng.normal(0, 0.035, (n, 3, 3))
    sdv = np.stack([np.where(rng.random(n) < 0.3, rng.choice([0.0, 1.0], n), rng.uniform(0, 1, n)),
                    rng.uniform(-100, 900, n), rng.uniform(-100, 900, n)])
    Fk = soa(Fs)
    sn, S, w = O.ref_sm160(MATS[mat], Fk, sdv)
    return [(mat, Fk, sdv, sn, S, w)]


def run_special(rng):
    """Identity, pure rotations, NaN input (eigen-solver bail-out), and a
    search for Newton iteration-cap cases (singular 2x2 Jacobians)."""
    Fs, mats, sdvs = [], [], []
    for m in range(len(MATS)):
        Fs.append(np.eye(3)); mats.append(m); sdvs.append([0.0, 0.0, 0.0])
        Fs.append(rot(rng)); mats.append(m); sdvs.append([0.3, 100.0, 50.0])
    nanF = np.eye(3); nanF[0, 1] = np.nan
    Fs.append(nanF); mats.append(0); sdvs.append([0.0, 0.0, 0.0])
    # Newton cap: search random histories whose 2x2 Newton matrix degenerates
    cand_F = np.eye(3) + rng.normal(0, 0.05, (20000, 3, 3))
    cand_sdv = np.stack([rng.uniform(0, 1, 20000), rng.uniform(-2000, 2000, 20000),
                         rng.uniform(-2000, 2000, 20000)])
    cand_mat = rng.integers(0, len(MATS), 20000).astype(np.int32)
    _, _, w = O.ref_sm160(MATS[cand_mat], soa(cand_F), cand_sdv)
    hits = np.where(w & 2)[0][:16]
    for h in hits:
        Fs.append(cand_F[h]); mats.append(int(cand_mat[h])); sdvs.append(list(cand_sdv[:, h]))
    mat = np.array(mats, dtype=np.int32)
    Fk = soa(np.array(Fs))
    sdv = np.array(sdvs).T.copy()
    sn, S, w = O.ref_sm160(MATS[mat], Fk, sdv)
    return [(mat, Fk, sdv, sn, S, w)], len(hits)


def gen_sm160():
    rng = np.random.default_rng(20261202)
    rows = run_paths(rng) + run_random(rng)
    special, nhits = run_special(rng)
    rows += special
    cat = lambda i, ax: np.concatenate([r[i] for r in rows], axis=ax)
    d = dict(mats=MATS, mat=cat(0, 0), F=cat(1, 1), sdv_old=cat(2, 1),
             sdv_new=cat(3, 1), S=cat(4, 1), warn=cat(5, 0))
    np.savez_compressed(os.path.join(OUT, "sm160_golden.npz"), **d)
    print("sm160 golden: n =", d["F"].shape[1], "newton-cap points:", nhits,
          "warn counts:", {b: int(np.sum(d["warn"] & b > 0)) for b in (1, 2, 4)})


def gen_hex8():
    rng = np.random.default_rng(8)
    n = 64
    ref = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                    [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=np.float64)
    elcod = np.zeros((n, 8, 3)); eldis = np.zeros((n, 8, 3))
    for e in range(n):
        sc = rng.uniform(0.2, 3.0, 3)
        A = np.diag(sc) + rng.normal(0, 0.15, (3, 3))
        elcod[e] = (ref @ A.T) * 0.5 + rng.normal(0, 0.06, (8, 3)) + rng.normal(0, 5, 3)
        eldis[e] = rng.normal(0, 0.02, (8, 3)) * sc.mean()
    r = O.ref_hex8(elcod, eldis)
    np.savez_compressed(os.path.join(OUT, "hex8_golden.npz"), elcod=elcod, eldis=eldis, **r)
    print("hex8 golden: n =", n)


if __name__ == "__main__":
    if not O.ref_available():
        sys.exit("oracle/_ref not built: run `make -C oracle ref` where the reference sources are")
    gen_sm160()
    gen_hex8()
