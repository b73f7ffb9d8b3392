"""The oracle's neo-Hookean law (sld_stress_model_102.f90, Ansys form),
pinned without the reference binary (its module chain needs MPI): the PK2
stress must be the derivative of the strain energy
W(C) = mu/2 (J^(-2/3) tr C - 3) + K/2 (J - 1)^2, vanish at F = I, reduce to
the pressure K J (J - 1) C^-1 under pure dilatation, and be frame
indifferent.  Planted mistakes (a dropped 1/3, a sign, J for J^(-2/3)) fail
the energy check."""
import numpy as np
import pytest

from oracle import oracle_py as O

MU, K = 0.03, 1.0            # tissue around the SMA tube, MPa


def energy(C, mu=MU, k=K):
    J = np.sqrt(np.linalg.det(C))
    return 0.5 * mu * (J ** (-2.0 / 3.0) * np.trace(C) - 3.0) + 0.5 * k * (J - 1.0) ** 2


def S_from_energy(C, h=1e-6):
    """S_ij = 2 dW/dC_ij with symmetric central differences."""
    S = np.zeros((3, 3))
    for i in range(3):
        for j in range(i, 3):
            E = np.zeros((3, 3))
            E[i, j] = E[j, i] = h if i != j else h
            d = (energy(C + E) - energy(C - E)) / (2 * h)
            S[i, j] = S[j, i] = (2 * d) if i == j else d
    return S


def test_zero_stress_at_identity():
    np.testing.assert_allclose(O.nh102_point([MU, K], np.eye(3)), 0.0, atol=1e-15)


@pytest.mark.parametrize("seed", range(6))
def test_pk2_is_the_energy_derivative(seed):
    rng = np.random.default_rng(seed)
    F = np.eye(3) + rng.normal(0, 0.15, (3, 3))
    assert np.linalg.det(F) > 0
    S = O.nh102_point([MU, K], F)
    Sfd = S_from_energy(F.T @ F)
    np.testing.assert_allclose(S, Sfd, rtol=1e-6, atol=1e-9 * np.abs(Sfd).max())


def test_pure_dilatation_is_pressure_only():
    for lam in (0.9, 1.0, 1.07, 1.3):
        S = O.nh102_point([MU, K], lam * np.eye(3))
        J = lam ** 3
        np.testing.assert_allclose(S, J * K * (J - 1) / lam ** 2 * np.eye(3), rtol=1e-13, atol=1e-16)


def test_frame_indifference():
    rng = np.random.default_rng(7)
    F = np.eye(3) + rng.normal(0, 0.2, (3, 3))
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(Q) < 0:
        Q[:, 0] *= -1
    np.testing.assert_allclose(O.nh102_point([MU, K], Q @ F), O.nh102_point([MU, K], F), rtol=1e-12,
                               atol=1e-15)


def test_uniaxial_incompressible_limit():
    """Isochoric stretch diag(l, l^-1/2, l^-1/2): Cauchy stress difference
    sigma_11 - sigma_22 = mu (l^2 - 1/l) (J = 1)."""
    lam = 1.25
    F = np.diag([lam, lam ** -0.5, lam ** -0.5])
    S = O.nh102_point([MU, K], F)
    sig = F @ S @ F.T
    assert abs((sig[0, 0] - sig[1, 1]) - MU * (lam ** 2 - 1 / lam)) < 1e-13
