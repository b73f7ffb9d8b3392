"""TEST INFRASTRUCTURE ONLY -- FP64 oracle of the MLP surrogate
(16 -> 128 x4 -> 12, SiLU), with the device's BF16 rounding points emulated
(inputs, hidden activations and the back-propagated pre-activation
gradients are rounded to BF16 round-to-nearest-even; everything else FP64).

Pinned by central finite differences of the FP64 loss (tests/test_oracle_mlp.py).
The reference has no surrogate; its per-Gauss-point ANN hook
(alya-ann/mod_ann_framework.f90:489-556) fixes only the call shape.
"""
import numpy as np

NIN, NH, NOUT = 16, 128, 12
SIZES = [(NH, NIN), (NH, NH), (NH, NH), (NH, NH), (NOUT, NH)]   # (n_out, n_in)


def unpack(W, b):
    """Canonical flat params -> lists of layer matrices [n_out][n_in], biases."""
    Ws, bs, o, ob = [], [], 0, 0
    for (n, k) in SIZES:
        Ws.append(np.asarray(W[o:o + n * k], dtype=np.float64).reshape(n, k))
        bs.append(np.asarray(b[ob:ob + n], dtype=np.float64))
        o += n * k
        ob += n
    return Ws, bs


def pack(Ws, bs):
    return (np.concatenate([w.ravel() for w in Ws]).astype(np.float32),
            np.concatenate(bs).astype(np.float32))


def bf16(x):
    """Round to BF16 (round to nearest even) and return as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(W, b, X, emulate=True):
    """Returns (Y [n][12], cache).  emulate: BF16 weights/activations."""
    Ws, bs = unpack(W, b)
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    x = r(X)
    xs, zs = [x], []
    for l, (w, bb) in enumerate(zip(Ws, bs)):
        z = x @ r(w).T + bb
        zs.append(z)
        if l < 4:
            x = r(silu(z))
            xs.append(x)
    return zs[-1], (Ws, bs, xs, zs)


def loss_and_grad(W, b, X, T, w=None, emulate=True):
    """L = (1/n) sum_m sum_j w_j (y - t)^2 and its gradient (canonical)."""
    n = X.shape[0]
    w = np.ones(NOUT) if w is None else np.asarray(w, dtype=np.float64)
    Y, (Ws, bs, xs, zs) = forward(W, b, X, emulate)
    d = Y - T
    L = float(np.sum(w * d * d) / n)
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    dz = 2.0 * w * d / n
    gWs, gbs = [None] * 5, [None] * 5
    for l in range(4, -1, -1):
        dzr = r(dz)
        gWs[l] = dzr.T @ xs[l]
        gbs[l] = dz.sum(axis=0)
        if l > 0:
            dy = dzr @ r(Ws[l])
            dz = dy * silu_grad(zs[l - 1])
    gW, gb = pack(gWs, gbs)
    return L, gW.astype(np.float64), gb.astype(np.float64)


def adam(W, b, gW, gb, state, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step in place; state = dict(mW, vW, mb, vb, t)."""
    state["t"] += 1
    t = state["t"]
    for p, g, m, v in ((W, gW, state["mW"], state["vW"]), (b, gb, state["mb"], state["vb"])):
        m[:] = beta1 * m + (1 - beta1) * g
        v[:] = beta2 * v + (1 - beta2) * g * g
        p -= lr * (m / (1 - beta1 ** t)) / (np.sqrt(v / (1 - beta2 ** t)) + eps)
