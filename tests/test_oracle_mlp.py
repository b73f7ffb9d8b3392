"""The FP64 MLP oracle: back-propagation pinned by central finite
differences of the FP64 loss; BF16 rounding helper pinned by exact cases."""
import numpy as np

from oracle import mlp_ref as M


def _params(seed=0):
    rng = np.random.default_rng(seed)
    Ws = [rng.uniform(-1, 1, (n, k)) * np.sqrt(6.0 / k) for (n, k) in M.SIZES]
    bs = [rng.normal(0, 0.1, n) for (n, _) in M.SIZES]
    W, b = M.pack(Ws, bs)
    return W.astype(np.float64), b.astype(np.float64)


def test_bf16_rounding_exact_cases():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 0.0], dtype=np.float32)
    r = M.bf16(x)
    # 1 + 2^-8 is the midpoint between 1 and 1 + 2^-7: ties to even -> 1
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -7
    assert r[3] == 1.0 + 2 ** -6 and r[4] == -2.5 and r[5] == 0.0


def test_gradient_matches_finite_differences():
    W, b = _params(1)
    rng = np.random.default_rng(2)
    X = rng.normal(size=(8, M.NIN))
    T = rng.normal(size=(8, M.NOUT))
    L, gW, gb = M.loss_and_grad(W, b, X, T, emulate=False)
    h = 1e-6
    for idx in rng.choice(W.size, 25, replace=False):
        Wp = W.copy(); Wp[idx] += h
        Wm = W.copy(); Wm[idx] -= h
        fd = (M.loss_and_grad(Wp, b, X, T, emulate=False)[0] - M.loss_and_grad(Wm, b, X, T, emulate=False)[0]) / (2 * h)
        assert abs(fd - gW[idx]) <= 1e-6 * max(1.0, abs(fd))
    for idx in range(0, b.size, 37):
        bp = b.copy(); bp[idx] += h
        bm = b.copy(); bm[idx] -= h
        fd = (M.loss_and_grad(W, bp, X, T, emulate=False)[0] - M.loss_and_grad(W, bm, X, T, emulate=False)[0]) / (2 * h)
        assert abs(fd - gb[idx]) <= 1e-6 * max(1.0, abs(fd))


def test_emulated_close_to_fp64():
    W, b = _params(3)
    X = np.random.default_rng(4).normal(size=(64, M.NIN))
    y1, _ = M.forward(W, b, X, emulate=True)
    y2, _ = M.forward(W, b, X, emulate=False)
    assert np.linalg.norm(y1 - y2) / np.linalg.norm(y2) < 3e-2


def test_adam_first_step_is_signed_lr():
    W = np.zeros(3); b = np.zeros(1)
    st = dict(mW=np.zeros(3), vW=np.zeros(3), mb=np.zeros(1), vb=np.zeros(1), t=0)
    M.adam(W, b, np.array([2.0, -3.0, 0.5]), np.array([1.0]), st, lr=0.01)
    np.testing.assert_allclose(W, [-0.01, 0.01, -0.01], rtol=1e-6)
