"""The surrogate prepass's closed-form principal deviatoric strains
(sma::principal_dev_fp32, the trigonometric roots of the characteristic
polynomial) against numpy's symmetric eigen-solver, the function compiled
for the host from the device header: generic, two- and three-fold
degenerate and tiny tensors, traces left by rounding."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc")

SRC = r'''
#include "sma_law.hpp"
#include <cstdio>
int main() {
  float s[6];
  while (scanf("%f %f %f %f %f %f", &s[0], &s[1], &s[2], &s[3], &s[4], &s[5]) == 6) {
    float l0, l1, l2;
    sma::principal_dev_fp32(s, l0, l1, l2);
    printf("%.9e %.9e %.9e\n", l0, l1, l2);
  }
  return 0;
}
'''


@pytest.fixture(scope="module")
def prog(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    d = tmp_path_factory.mktemp("pdev")
    src = d / "pdev.cpp"
    src.write_text(SRC)
    exe = d / "pdev"
    subprocess.run([HIPCC, "-x", "hip", "--offload-host-only", "-O2", "-std=c++17",
                    "-I" + os.path.join(ROOT, "alya_sma_amd", "csrc"), str(src), "-o", str(exe)],
                   check=True, capture_output=True)
    return str(exe)


def _run(exe, S):
    """S: [n, 3, 3] symmetric -> [n, 3] from the compiled function."""
    s6 = np.stack([S[:, 0, 0], S[:, 0, 1], S[:, 0, 2], S[:, 1, 1], S[:, 1, 2], S[:, 2, 2]], axis=1)
    s6 = s6.astype(np.float32)
    text = "\n".join(" ".join(f"{v:.9e}" for v in row) for row in s6) + "\n"
    out = subprocess.run([exe], input=text, capture_output=True, text=True, check=True).stdout
    return np.array([[float(x) for x in line.split()] for line in out.splitlines()]), s6


def _dev(S):
    return S - np.trace(S, axis1=1, axis2=2)[:, None, None] / 3.0 * np.eye(3)


def _reference(s6):
    S = np.zeros((len(s6), 3, 3))
    idx = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    for k, (i, j) in enumerate(idx):
        S[:, i, j] = S[:, j, i] = s6[:, k].astype(np.float64)
    return np.linalg.eigvalsh(_dev(S))[:, ::-1]


def test_generic_tensors(prog):
    rng = np.random.default_rng(0)
    A = rng.normal(size=(400, 3, 3))
    S = _dev(A + A.transpose(0, 2, 1))
    S *= (10.0 ** rng.uniform(-9, 0, size=400))[:, None, None]
    got, s6 = _run(prog, S)
    ref = _reference(s6)
    scale = np.linalg.norm(ref, axis=1)
    assert np.all(got[:, 0] >= got[:, 1]) and np.all(got[:, 1] >= got[:, 2])
    assert np.all(np.abs(got - ref) <= 2e-6 * scale[:, None]), np.max(np.abs(got - ref) / scale[:, None])
    # traceless, as the deviator's principal values are
    assert np.all(np.abs(got.sum(axis=1)) <= 1e-6 * scale)


def test_degenerate_and_rotated(prog):
    rng = np.random.default_rng(1)
    diag = [np.diag([2.0, -1.0, -1.0]), np.diag([1.0, 1.0, -2.0]), np.diag([1.0, 0.0, -1.0]),
            np.diag([1.0 + 1e-4, -0.5, -0.5 - 1e-4])]
    S = []
    for D in diag:
        for _ in range(25):
            Q, _r = np.linalg.qr(rng.normal(size=(3, 3)))
            S.append(0.03 * Q @ D @ Q.T)
    S = np.array(S)
    got, s6 = _run(prog, S)
    ref = _reference(s6)
    scale = np.linalg.norm(ref, axis=1)
    # a double root costs the trigonometric form about sqrt(FP32 eps) of the
    # scale in the pair's split; the network's inputs are rounded to BF16
    # (2^-8 relative) before use
    assert np.all(np.abs(got - ref) <= 5e-4 * scale[:, None]), np.max(np.abs(got - ref) / scale[:, None])


def test_tiny_and_trace(prog):
    S = np.array([np.zeros((3, 3)), 1e-14 * np.eye(3), np.diag([1e-13, -1e-13, 0.0]),
                  np.diag([0.011, 0.001, 0.001])])   # the last carries a trace
    got, s6 = _run(prog, S)
    np.testing.assert_array_equal(got[:3], 0.0)
    ref = _reference(s6)
    assert np.all(np.abs(got[3] - ref[3]) <= 2e-6 * np.linalg.norm(ref[3]))
