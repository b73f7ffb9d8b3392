"""Synthetic problem builders: sizes, seeding and independence of the
batched specimens (the union mesh of uneven SMA bars)."""
import numpy as np

from alya_sma_amd import problems


def test_batched_specimens_sizes_and_seed():
    p = problems.batched_specimens(n=32, median=300, sigma=0.7, lo=100, hi=2000, nsteps=100)
    s = p["specimens"]
    assert len(s["elements"]) == 32 and p["props"].shape == (32, 18)
    # cubic elements: n_e = nx * ny^2 within rounding of the drawn size
    assert np.all(s["elements"] >= 0.5 * 100) and np.all(s["elements"] <= 1.5 * 2000)
    assert p["lnods"].shape[0] == s["elements"].sum()
    assert np.all((s["amplitude"] >= 0.03) & (s["amplitude"] <= 0.07))
    assert np.all(np.abs(p["props"][:, 0] / problems.NITINOL[0] - 1) <= 0.1 + 1e-12)
    q = problems.batched_specimens(n=32, median=300, sigma=0.7, lo=100, hi=2000, nsteps=100)
    assert np.array_equal(p["lnods"], q["lnods"]) and np.array_equal(p["func_table"], q["func_table"])


def test_batched_specimens_are_disjoint_with_own_material_and_load():
    p = problems.batched_specimens(n=5, median=60, lo=20, hi=200, nsteps=100)
    mat, ln = p["mat_of_elem"], p["lnods"]
    node_spec = np.full(p["coord"].shape[0], -1)
    for s in range(5):
        nodes = np.unique(ln[mat == s])
        assert np.all(node_spec[nodes] == -1)          # no node shared between specimens
        node_spec[nodes] = s
        tip = nodes[p["func_of_node"][nodes] >= 0]
        assert np.all(p["func_of_node"][tip] == s)       # each loaded end follows its own function
        assert np.all(p["fixmask"][tip] & 1)
    assert np.all(node_spec >= 0)
    # one step size for all: the smallest critical step times the safety factor
    h = 1.0 / p["specimens"]["dims"][:, 1]
    E, nu, rho = p["props"][:, 0], p["props"][:, 1], p["rho"]
    c = np.sqrt((nu * E / ((1 + nu) * (1 - 2 * nu)) + E / (1 + nu)) / rho)
    assert np.isclose(p["run"]["dt"], 0.5 * np.min(h / c), rtol=1e-12)
