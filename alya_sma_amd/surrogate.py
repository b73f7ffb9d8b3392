"""Host driver of the MLP surrogate of the SMA return map (C ABI
sma_mlp_* in include/alya_sma.h): data generation from the exact FP64
return map on the GPU, training with the project's BF16 MFMA forward /
backward / Adam kernels, inference, save/load.

Device memory comes from the library's own allocator (alya_sma_amd.device);
every computation runs in libalya_sma_amd.so.
"""
import argparse
import ctypes as C
import math
import os
import time

import numpy as np

from . import problems
from ._lib import check, lib
from .device import DeviceArray, synchronize

NIN, NOUT = 16, 12
NWEIGHT, NBIAS = 52736, 524
# loss weights per output: principal stresses, transformation strains, xS,
# F_AS, F_SA, branch flags, xS increment.  The step uses only the increment
# (stress and loading values are recomputed exactly from it), so it carries
# the weight; F_AS/F_SA are left out: stored at the pre-clamp xS they have
# heavy tails (std ~ 23 against ~0.3 for the other targets) that swamp the
# gradient.
LOSS_WEIGHTS = np.array([1, 1, 1, 1, 1, 1, 1, 0, 0, 0.25, 0.25, 100], dtype=np.float32)


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class Surrogate:
    def __init__(self, seed=0):
        h = C.c_void_p()
        check(lib().sma_mlp_create(int(seed), C.byref(h)), "sma_mlp_create")
        self.h = h
        self.set_loss_weights(LOSS_WEIGHTS)

    def close(self):
        if getattr(self, "h", None):
            lib().sma_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ----------------------------------------------------------- parameters
    def get_params(self):
        W = np.zeros(NWEIGHT, np.float32)
        b = np.zeros(NBIAS, np.float32)
        check(lib().sma_mlp_get_params(self.h, _ptr(W), _ptr(b)), "sma_mlp_get_params")
        return W, b

    def set_params(self, W, b):
        W = np.ascontiguousarray(W, np.float32)
        b = np.ascontiguousarray(b, np.float32)
        assert W.size == NWEIGHT and b.size == NBIAS
        check(lib().sma_mlp_set_params(self.h, _ptr(W), _ptr(b)), "sma_mlp_set_params")

    def set_loss_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        check(lib().sma_mlp_set_loss_weights(self.h, _ptr(w)), "sma_mlp_set_loss_weights")

    def save(self, path):
        W, b = self.get_params()
        np.savez_compressed(path, W=W, b=b)

    def load(self, path):
        d = np.load(path)
        self.set_params(d["W"], d["b"])

    # -------------------------------------------------------------- compute
    def forward(self, X, n=None):
        """X: DeviceArray [n, 16] float32 -> DeviceArray [n, 12]."""
        n = X.shape[0] if n is None else n
        Y = DeviceArray((n, NOUT), np.float32)
        check(lib().sma_mlp_forward_device(self.h, n, X.ptr, Y.ptr, None), "sma_mlp_forward_device")
        return Y

    def grad(self, X, T, n=None, x_ptr=None, t_ptr=None):
        n = X.shape[0] if n is None else n
        gW = DeviceArray(NWEIGHT, np.float32)
        gb = DeviceArray(NBIAS, np.float32)
        loss = C.c_float()
        check(lib().sma_mlp_grad_device(self.h, n, x_ptr or X.ptr, t_ptr or T.ptr, gW.ptr, gb.ptr,
                                        C.byref(loss), None), "sma_mlp_grad_device")
        return loss.value, gW.to_host(), gb.to_host()

    def train_step(self, x_ptr, t_ptr, n, lr, want_loss=False):
        loss = C.c_float()
        check(lib().sma_mlp_train_step_device(self.h, int(n), x_ptr, t_ptr, float(lr),
                                              C.byref(loss) if want_loss else None, None),
              "sma_mlp_train_step_device")
        return loss.value if want_loss else None

    # ----------------------------------------------------------------- data
    @staticmethod
    def generate(n, seed=2, props=problems.NITINOL, strain_sd=0.02, step_sd=0.003, jitter=0.1):
        """Config-2 samples labelled by the exact FP64 return map on the GPU.
        Returns DeviceArrays X [n,16], T [n,12], flags [n]."""
        p = DeviceArray.from_host(np.asarray(props, np.float64))
        X = DeviceArray((n, NIN), np.float32)
        T = DeviceArray((n, NOUT), np.float32)
        F = DeviceArray(n, np.uint32)
        check(lib().sma_mlp_gen_samples_device(n, int(seed), p.ptr, float(strain_sd), float(step_sd),
                                               float(jitter), X.ptr, T.ptr, F.ptr, None),
              "sma_mlp_gen_samples_device")
        synchronize()
        p.free()
        return X, T, F

    def fit(self, X, T, steps=4000, batch=16384, lr=2e-3, lr_min=1e-5, log_every=500, log=print):
        """Adam with cosine decay over consecutive mini-batches of the iid
        samples (X, T), wrapping around at the end."""
        n = X.shape[0]
        batch = min(batch, n)
        pos = 0
        hist = []
        t0 = time.perf_counter()
        for k in range(steps):
            if pos + batch > n:
                pos = 0
            cur = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * k / max(1, steps - 1)))
            want = (k % log_every == 0) or k == steps - 1
            loss = self.train_step(X.rows(pos, pos + batch), T.rows(pos, pos + batch), batch, cur,
                                   want_loss=want)
            pos += batch
            if want:
                hist.append((k, loss))
                log(f"surrogate step {k}: loss {loss:.3e} lr {cur:.2e} "
                    f"({time.perf_counter() - t0:.1f} s)")
        synchronize()
        return hist


# Network trained by train_default() (python -m alya_sma_amd.surrogate).
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "sma160_mlp_nitinol.npz")


def train_default(seed=1, n_samples=20_000_000, steps=20000, batch=16384, lr=1e-3, log=print):
    """The configuration-2 network: n_samples labelled by the exact return
    map (NITINOL +-10%), Adam with cosine decay over `steps` mini-batches."""
    X, T, F = Surrogate.generate(n_samples, seed=21)
    s = Surrogate(seed=seed)
    s.fit(X, T, steps=steps, batch=batch, lr=lr, lr_min=lr * 1e-2, log_every=max(1, steps // 5), log=log)
    for a in (X, T, F):
        a.free()
    return s


def load_default():
    """The committed network, or None when the file is absent."""
    if not os.path.exists(WEIGHTS):
        return None
    s = Surrogate()
    s.load(WEIGHTS)
    return s


def transformation_error(sur, X, T, flags):
    """Accuracy of the trained network on transforming points: absolute error
    of the xS update it proposes, and of xS itself."""
    Y = sur.forward(X).to_host()
    Th, fl = T.to_host(), flags.to_host()
    m = (fl & 3) != 0
    dx_err = np.abs(Y[:, 11] - Th[:, 11])[m]
    return {"transforming_fraction": float(m.mean()),
            "dx_abs_err_median": float(np.median(dx_err)),
            "dx_abs_err_p99": float(np.quantile(dx_err, 0.99)),
            "xs_abs_err_median": float(np.median(np.abs(Y[:, 6] - Th[:, 6])[m]))}


def main():
    ap = argparse.ArgumentParser(description="train the SMA return-map surrogate on the GPU")
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--samples", type=int, default=20_000_000)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--seed", type=int, default=1)
    a = ap.parse_args()
    t0 = time.perf_counter()
    s = train_default(seed=a.seed, n_samples=a.samples, steps=a.steps)
    Xv, Tv, Fv = Surrogate.generate(200_000, seed=99)
    print(f"trained in {time.perf_counter() - t0:.1f} s; holdout {transformation_error(s, Xv, Tv, Fv)}",
          flush=True)
    os.makedirs(os.path.dirname(a.out), exist_ok=True)
    s.save(a.out)


if __name__ == "__main__":
    main()
