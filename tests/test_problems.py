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


def test_cylinder_problem_two_materials():
    p = problems.cylinder_problem(ntheta=12, nz=3, nr_tube=2, nr_tissue=3, nsteps=100)
    ln, mat = p["lnods"], p["mat_of_elem"]
    assert ln.shape[0] == 5 * 12 * 3
    assert np.array_equal(p["law_of_mat"], [160, 102])
    # tube elements lie inside r = 4.3, tissue elements outside
    rc = np.linalg.norm(p["coord"][ln][:, :, :2].mean(axis=1), axis=1)
    assert np.all(rc[mat == 0] < 4.3) and np.all(rc[mat == 1] > 4.3)
    # the inner surface follows one radial function per station, u_z = 0 on z = 0
    r = np.linalg.norm(p["coord"][:, :2], axis=1)
    inner = np.abs(r - 4.0) < 1e-9
    assert np.all(p["fixmask"][inner] & 3 == 3) and np.all(p["func_of_node"][inner] >= 0)
    f = p["func_of_node"][inner]
    d = p["func_table"][f, 1, :2]                       # crimp direction at the peak
    x = p["coord"][inner, :2] / 4.0
    np.testing.assert_allclose(d, -0.4 * x, atol=1e-12)
    assert np.all(p["fixmask"][np.abs(p["coord"][:, 2]) < 1e-9] & 4)
