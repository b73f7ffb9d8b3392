"""The algorithmic FLOP counts bench.py uses are the ones the instrumented
CPU oracle prints (oracle/flopcount.cpp: the oracle compiled with a
counting FP64 type), and the frozen table oracle/flops_per_gp.txt is its
current output."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def _table(text):
    out = {}
    for line in text.splitlines():
        if line.startswith("#") or not line.strip():
            continue
        name, val = line.split()[:2]
        out[name] = float(val)
    return out


def test_frozen_counts_match_the_instrumented_oracle():
    odir = os.path.join(ROOT, "oracle")
    subprocess.run(["make", "-s", "-C", odir, "flopcount"], check=True)
    run = subprocess.run([os.path.join(odir, "_build", "flopcount")], check=True,
                         capture_output=True, text=True).stdout
    with open(os.path.join(odir, "flops_per_gp.txt")) as f:
        frozen = f.read()
    assert _table(run) == _table(frozen)
    t = _table(run)
    for k, v in bench.FLOPS_PER_GP.items():
        assert t[k + "_per_gp"] == v, k
    assert t["nodal_per_node"] == bench.FLOPS_PER_NODE
