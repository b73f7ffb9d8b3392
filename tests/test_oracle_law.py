"""The CPU oracle of the SMA law against the reference's own compiled law
(golden vectors, tests/golden/gen_golden.py) and against what the mechanics
fixes independently of both: the Hencky closed form, the saturated branch,
frame indifference, and the branch inequalities of
sld_stress_model_160.f90:351,368."""
import numpy as np
import pytest

from conftest import soa, unsoa
from oracle import oracle_py as O

NIT = np.array([60000, 0.33, 0.05, 400, 500, 0, 0, 0, 0, 250, 150, 0, 0, 0, 0, 1, 1, 0.0])


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def test_golden_bitwise(golden_sm160):
    g = golden_sm160
    sn, S, fl = O.sm160_eval(g["mats"], g["F"], g["sdv_old"], mat_of_gp=g["mat"])
    ok = np.isfinite(g["S"]).all(axis=0)
    # the oracle follows the reference operation by operation: equal bits
    assert np.array_equal(S[:, ok], g["S"][:, ok])
    assert np.array_equal(sn[:, ok], g["sdv_new"][:, ok])
    # NaN input: the reference returns NaN stress; so does the oracle
    assert np.all(np.isnan(S[:, ~ok]).any(axis=0))


def test_golden_warnings(golden_sm160):
    g = golden_sm160
    _, _, fl = O.sm160_eval(g["mats"], g["F"], g["sdv_old"], mat_of_gp=g["mat"])
    w = g["warn"]
    assert np.array_equal((w & 2) > 0, (fl & O.FLAG_NEWTON_FAIL) > 0)
    assert np.array_equal((w & 4) > 0, (fl & O.FLAG_EIGEN_FAIL) > 0)
    both = (fl & 3) == 3
    assert np.array_equal((w & 1) > 0, both)
    assert (w & 2).sum() > 0 and (w & 4).sum() > 0 and (w & 1).sum() > 0


def test_golden_covers_branches(golden_sm160):
    g = golden_sm160
    _, _, fl = O.sm160_eval(g["mats"], g["F"], g["sdv_old"], mat_of_gp=g["mat"])
    xs = g["sdv_new"][0]
    assert np.sum(fl & O.FLAG_KF_AS) > 100 and np.sum(fl & O.FLAG_KF_SA) > 100
    assert np.sum(xs == 1.0) > 10 and np.sum((xs > 0) & (xs < 1)) > 100
    assert np.sum((fl & 3) == 0) > 100


def _hencky_tau(F, E, nu):
    K = E / (3 - 6 * nu)
    G = E / (2 + 2 * nu)
    b = F @ F.T
    w, V = np.linalg.eigh(b)
    lnV = V @ np.diag(0.5 * np.log(w)) @ V.T
    th = np.trace(lnV)
    return K * th * np.eye(3) + 2 * G * (lnV - th / 3 * np.eye(3))


@pytest.mark.parametrize("kind", ["uniaxial", "shear", "hydrostatic", "general"])
def test_elastic_hencky_closed_form(kind):
    rng = np.random.default_rng(3)
    if kind == "uniaxial":
        F = np.diag([1.004, 0.9987, 0.9987])
    elif kind == "shear":
        F = np.eye(3); F[0, 1] = 0.006
    elif kind == "hydrostatic":
        F = 1.003 * np.eye(3)
    else:
        F = np.eye(3) + rng.normal(0, 0.002, (3, 3))
    sn, S, tau, fl = O.sm160_point(NIT, F, np.zeros(3))
    assert fl & 3 == 0 and sn[0] == 0.0            # below threshold: no transformation
    ref = _hencky_tau(F, 60000, 0.33)
    assert np.max(np.abs(tau - ref)) <= 1e-9 * np.max(np.abs(ref))
    Finv = np.linalg.inv(F)
    assert np.max(np.abs(S - Finv @ ref @ Finv.T)) <= 1e-9 * np.max(np.abs(ref))


def test_saturated_branch_closed_form():
    # xS_0 = 1 with loading: kf_AS needs xS_0 < 1, so xS stays 1 and
    # tau = K theta I + 2G (e - eL d) (sld_stress_model_160.f90:487-491)
    F = np.diag([1.08, 0.975, 0.975])
    sn, S, tau, fl = O.sm160_point(NIT, F, np.array([1.0, 0.0, 1e9]))
    assert sn[0] == 1.0
    K, G, eL = 60000 / (3 - 6 * .33), 60000 / (2 + 2 * .33), 0.05
    lam = np.log(np.diag(F))
    th = lam.sum()
    e = lam - th / 3
    d = e / np.linalg.norm(e)
    ref = np.diag(K * th + 2 * G * (e - eL * d))
    assert np.max(np.abs(tau - ref)) <= 1e-9 * np.max(np.abs(ref))


def test_frame_indifference():
    rng = np.random.default_rng(11)
    for _ in range(20):
        F = np.eye(3) + rng.normal(0, 0.03, (3, 3))
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        if np.linalg.det(q) < 0:
            q[:, 0] *= -1
        sdv = np.array([rng.uniform(0, 1), rng.uniform(0, 600), rng.uniform(0, 600)])
        s1, _, t1, f1 = O.sm160_point(NIT, F, sdv)
        s2, _, t2, f2 = O.sm160_point(NIT, q @ F, sdv)
        assert (f1 & 3) == (f2 & 3)
        np.testing.assert_allclose(s2, s1, rtol=1e-9, atol=1e-9)
        np.testing.assert_allclose(t2, q @ t1 @ q.T, rtol=0, atol=1e-8 * np.abs(t1).max())


def test_branch_inequalities():
    # monotone uniaxial loading from rest: AS starts once F_AS > Rs_AS
    Rs_AS = 400 * np.sqrt(2 / 3)
    G = 60000 / 2.66
    sdv = np.zeros(3)
    seen_as = False
    for k, eps in enumerate(np.linspace(0, 0.05, 60)):
        F = np.diag([np.exp(eps), np.exp(-0.33 * eps), np.exp(-0.33 * eps)])
        sn, _, _, fl = O.sm160_point(NIT, F, sdv)
        # F_AS at xS_0 = 2G e_n (alpha = 0, T inert)
        lam = np.log(np.diag(F)); e = lam - lam.sum() / 3
        FAS = 2 * G * np.linalg.norm(e)
        expect = FAS > sdv[1] and FAS - Rs_AS > 0 and sdv[0] < 1
        assert bool(fl & O.FLAG_KF_AS) == expect
        seen_as |= expect
        sdv = sn
    assert seen_as
    # unloading from a transformed state: SA once F_SA < F_SA_0 and F_SA < Rs_SA
    sn, _, _, fl = O.sm160_point(NIT, np.diag([1.004, 0.999, 0.999]), np.array([0.6, 600.0, 600.0]))
    assert fl & O.FLAG_KF_SA and not fl & O.FLAG_KF_AS and sn[0] < 0.6


def test_newton_cap_reverts_history(golden_sm160):
    g = golden_sm160
    idx = np.where(g["warn"] & 2)[0]
    assert len(idx) > 0
    sn, _, fl = O.sm160_eval(g["mats"], g["F"][:, idx], g["sdv_old"][:, idx], mat_of_gp=g["mat"][idx])
    assert np.all(fl & O.FLAG_NEWTON_FAIL)
    assert ((fl >> O.ITER_SHIFT) & 0xff == 65).all()
    # xS and both loading values revert to TIME_N (sld_stress_model_160.f90:459-465)
    xs0 = np.clip(g["sdv_old"][0, idx], 0, 1)
    np.testing.assert_array_equal(sn[0], xs0)
    np.testing.assert_array_equal(sn[1:], g["sdv_old"][1:, idx])


def test_single_newton_iteration_on_linear_branches(golden_sm160):
    # the residual is linear in dh when one branch is active, so Newton stops
    # after one update (the running-total quirk never matters there)
    g = golden_sm160
    _, _, fl = O.sm160_eval(g["mats"], g["F"], g["sdv_old"], mat_of_gp=g["mat"])
    one = ((fl & 3) == 1) | ((fl & 3) == 2)
    it = (fl >> O.ITER_SHIFT) & 0xff
    assert np.mean(it[one] <= 1) > 0.99


def test_eigen_sort_descending():
    rng = np.random.default_rng(5)
    for _ in range(50):
        A = rng.normal(size=(3, 3)); A = A @ A.T + np.eye(3)
        rc, w, Q = O.dsyevq3(A)
        assert rc == 0
        assert w[0] >= w[1] >= w[2]
        np.testing.assert_allclose(Q @ np.diag(w) @ Q.T, A, atol=1e-12 * np.abs(A).max())
        np.testing.assert_allclose(np.sort(w)[::-1], np.linalg.eigvalsh(A)[::-1], rtol=1e-12)
    # repeated eigenvalues (exact uniaxial stretch): sorted, orthonormal basis
    rc, w, Q = O.dsyevq3(np.diag([1.2, 0.9, 0.9]))
    assert rc == 0 and w[0] == 1.2 and w[1] == w[2] == 0.9
    np.testing.assert_allclose(Q.T @ Q, np.eye(3), atol=1e-15)


def test_golden_paths_match_sequential_oracle(golden_sm160):
    # the first path step of every path starts from zero history
    g = golden_sm160
    assert np.all(g["sdv_old"][:, :96] == 0)
    # planted mistake: a sign flip in F_SA_0 must be caught by the goldens
    so = g["sdv_old"].copy(); so[2] = -so[2]
    sn, S, _ = O.sm160_eval(g["mats"], g["F"], so, mat_of_gp=g["mat"])
    ok = np.isfinite(g["S"]).all(axis=0)
    assert not np.array_equal(sn[:, ok], g["sdv_new"][:, ok])


def test_matches_ref_binary_when_present():
    if not O.ref_available():
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    rng = np.random.default_rng(99)
    n = 500
    Fs = np.eye(3)[None] + rng.normal(0, 0.03, (n, 3, 3))
    so = np.stack([rng.uniform(0, 1, n), rng.uniform(0, 800, n), rng.uniform(0, 800, n)])
    sn_r, S_r, _ = O.ref_sm160(np.tile(NIT, (n, 1)), soa(Fs), so)
    sn_o, S_o, _ = O.sm160_eval(NIT[None], soa(Fs), so)
    assert np.array_equal(S_r, S_o) and np.array_equal(sn_r, sn_o)


@pytest.mark.parametrize("scale", [1e-60, 1e-90, 1e-130, 1e-158, 1e-175])
def test_underflow_band_matches_ref_binary(scale):
    """Deformation gradients within about 1e-162..1e-77 of the identity (the
    displacement tails ahead of a loading wave pass through them): the
    reference's eigen-solver (dsytrd3/dsyevq3) forms fourth powers of the
    off-diagonals of b, they underflow, and its law returns NaN stress and
    loading functions there; at 1e-60 and at 1e-175 (the off-diagonal
    products of b vanish) it returns finite values.  The oracle restates that bit for bit, NaNs included,
    which is why the full-size parity tests compare states the loading wave
    has crossed."""
    if not O.ref_available():
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    rng = np.random.default_rng(7)
    n = 64
    Fs = np.eye(3)[None] + scale * rng.normal(size=(n, 3, 3))
    so = np.zeros((3, n))
    sn_r, S_r, _ = O.ref_sm160(np.tile(NIT, (n, 1)), soa(Fs), so)
    sn_o, S_o, _ = O.sm160_eval(NIT[None], soa(Fs), so)
    nan_r = ~np.isfinite(S_r)
    assert np.array_equal(nan_r, ~np.isfinite(S_o))
    assert np.array_equal(S_r[~nan_r], S_o[~nan_r])
    if 1e-162 < scale < 1e-77:
        assert nan_r.any(axis=0).all()
    else:
        assert not nan_r.any()
