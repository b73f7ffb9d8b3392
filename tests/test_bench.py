"""Host logic of bench.py (no GPU): the one-GPU guard, the element kernel's
algorithmic bytes, and the traffic figures read from the committed rocprofv3
summary."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402
from alya_sma_amd import problems  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stdout


def test_algorithmic_bytes():
    # per GP: 48 history + 4 connectivity + 1 per-element modulus; per node
    # 48 read + the internal force written (3 x 8 or 3 x 4)
    prob = problems.cantilever_problem(nx=10, ny=2, nz=2, lx=10.0, ly=2.0, lz=2.0)
    ne, npn = 40, 11 * 3 * 3
    assert bench.algorithmic_bytes(prob, "fp64") == ne * 8 * 53 + npn * 72
    assert bench.algorithmic_bytes(prob, "fp32") == ne * 8 * 53 + npn * 60
    # the design's gather layout: 2 B moduli, 4 B slot and the corner force
    assert bench.design_bytes(prob, "fp32", "gather") == ne * 8 * 70 + npn * 48
    assert bench.design_bytes(prob, "fp64", "gather") == ne * 8 * 82 + npn * 48


def test_algorithmic_bytes_without_history_on_neo_hookean_elements():
    prob = problems.cylinder_problem(ntheta=8, nz=2, nr_tube=1, nr_tissue=2)
    ne, npn = prob["lnods"].shape[0], prob["coord"].shape[0]
    n_sma = int(np.count_nonzero(prob["mat_of_elem"] == 0))
    assert prob.get("E_of_elem") is None
    assert bench.algorithmic_bytes(prob, "fp32") == 8 * (n_sma * 52 + (ne - n_sma) * 4) + npn * 60


def test_algorithmic_flops_mix():
    prob = problems.cylinder_problem(ntheta=8, nz=2, nr_tube=1, nr_tissue=2)
    ne = prob["lnods"].shape[0]
    n_sma = 8 * int(np.count_nonzero(prob["mat_of_elem"] == 0))
    f = bench.FLOPS_PER_GP
    base = 8 * ne * (f["kinematics"] + f["force"]) + n_sma * f["law_elastic"] \
        + (8 * ne - n_sma) * f["law_nh102"]
    assert bench.algorithmic_flops(prob, 0, 0) == base
    assert bench.algorithmic_flops(prob, 10, 4) == \
        base + 10 * (f["law_AS"] - f["law_elastic"]) + 4 * (f["law_SA"] - f["law_elastic"])


def test_kernel_names_and_committed_traffic():
    cfg = {"assembly": "gather", "elems_per_block": 8}
    k32 = bench.kernel_name("fp32", cfg)
    assert k32 == "sld_element<float, 8, false, 0, true, 1, false>"
    # each element's own law: a mixed mesh scattered; gathered (surrogate
    # mode too) the SMA range runs the SMA-only kernel
    assert bench.kernel_name("surrogate", cfg, mixed=True) == \
        "sld_element<float, 8, false, 2, true, 1, false>"
    assert bench.kernel_name("fp64", {"assembly": "scatter", "elems_per_block": 16}, mixed=True) == \
        "sld_element<double, 16, false, 0, false, 0, false>"
    assert bench.kernel_name("fp64", cfg, mixed=True).endswith(", 1, false>")
    t = bench.measured_traffic("cantilever", k32)
    assert t is not None and 0.9 * 98.09e6 < t < 2 * 98.09e6
    # the tube in tissue: the SMA launch and the neo-Hookean-only launch add
    lines = {}
    for line in open(bench.TRAFFIC_FILE):
        if line.startswith("traffic_bytes_per_launch cylinder sld_element<float"):
            name, val = line.split(None, 2)[2].rsplit("=", 1)
            lines[name.strip()] = float(val)
    assert len(lines) == 2
    assert bench.measured_traffic("cylinder", k32) == sum(lines.values())
    assert bench.measured_traffic("no such workload", k32) is None
