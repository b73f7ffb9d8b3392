"""Alya deck reader on the reference's own uniaxial test deck
(tests/uniaxial_test/sm154_d1C.*, kept as data under tests/golden/):
parsed geometry, boundary codes, CODES, NODES rules, functions, material
line and integrator; then the whole deck run by the CPU oracle shows the
superelastic cycle its load table drives."""
import os

import numpy as np

from alya_sma_amd import deck

HERE = os.path.dirname(os.path.abspath(__file__))
CASE = os.path.join(HERE, "golden", "uniaxial_test", "sm154_d1C.dat")


def test_deck_parses_the_uniaxial_case():
    p = deck.read_case(CASE)
    assert p["coord"].shape == (8, 3) and p["lnods"].tolist() == [[0, 1, 2, 3, 4, 5, 6, 7]]
    # node codes from the six boundary faces (fix.dat: boundary i has code i)
    assert p["node_codes"][0] == {1, 3, 6} and p["node_codes"][6] == {2, 4, 5}
    # CODES, NODES: "1 & 3 & 4 111 ... MODULE_FUNCTION= 1" (node 2) etc.
    assert p["fixmask"].tolist() == [7, 7, 5, 5, 3, 3, 1, 1]
    assert p["func_of_node"].tolist() == [-1, 0, 0, -1, -1, 0, 0, -1]
    tab = p["func_table"][0]
    np.testing.assert_allclose(tab[:, 3], [0, 5e-3, 1e-2, 1.5e-2, 2e-2])
    np.testing.assert_allclose(tab[:, 0], [0, 0.1, 0, 0.04, 0])
    assert p["func_smooth"].tolist() == [1] and p["func_npts"].tolist() == [5]
    np.testing.assert_allclose(p["props"][0, :17], [1000, 0.3, 0.1, 90, 150, 0, 0, 0, 700, 70, 30,
                                                    0, 0, 0, 0, 10, 10])
    assert p["rho"].tolist() == [6.4e-9] and p["law_of_mat"].tolist() == [160]
    r = p["run"]
    assert (r["scheme"], r["phi"], r["safety"], r["t_end"]) == (1, 1.033, 0.1, 0.02)
    assert {k: v.tolist() for k, v in p["node_sets"].items()} == {1: [3], 2: [7]}


def test_deck_run_by_the_oracle_cycles_through_transformation():
    from alya_sma_amd import _lib
    from oracle import oracle_py as O
    p = deck.read_case(CASE)
    o = O.StepOracle(p)
    dt = p["run"]["safety"] * o.critical_dt()
    n = int(np.ceil(p["run"]["t_end"] / dt))
    xs, sx = [], []
    for k in range(40):
        o.step(dt, scheme=1, phi=p["run"]["phi"], n=n // 40)
        xs.append(o.sdv[0, ..., 0].mean())
        sx.append(o.S[..., 0].mean())
    xs, sx = np.array(xs), np.array(sx)
    # E = 1000 MPa, Rs_AS = 90 sqrt(2/3) = 73.5 MPa, Rf_AS = 122.5 MPa: the
    # first load (10% strain) passes the start of forward transformation but
    # not its finish; unloading and the 4% reload leave no martensite
    assert xs[:10].max() > 0.02 and np.argmax(xs[:12] > 0) >= 6
    assert 73.0 < sx[:12].max() < 122.5
    assert xs[-1] == 0.0 and abs(sx[-1]) < 1.0


def test_deck_neumann_keywords(tmp_path):
    """A deck of our own with CODES, BOUNDARIES (fixbo 3 + MODULE_FUNCTION),
    FOLLOWER_LOADS: ON under TRANSIENT conditions and FORCE: FIELD: the
    faces, codes and values read are those written, and the nodal-force
    ramp equals cutim / (t1 - t0) (sld_matrix.f90:249-262)."""
    from _decks import LZ, PRESSURE, T1, TIP_FORCE, write_beam_deck
    from oracle import oracle_py as O
    path, ref = write_beam_deck(str(tmp_path))
    p = deck.read_case(path)
    f = p["faces"]
    top = ref["code"] == 2
    np.testing.assert_array_equal(f["lnodb"], ref["faces"][top])
    assert f["code"].tolist() == [3] * int(top.sum())
    assert f["func"].tolist() == [0] * int(top.sum())
    assert f["transient"] and f["follower"]
    # the clamped end from the face codes, nothing else fixed
    clamped = np.abs(ref["coord"][:, 0]) < 1e-9
    assert (p["fixmask"][clamped] == 7).all() and (p["fixmask"][~clamped] == 0).all()
    fe = np.zeros_like(ref["coord"])
    fe[ref["tip"]] = TIP_FORCE
    np.testing.assert_array_equal(p["fext0"], fe)
    o = O.StepOracle(p)
    for t in (0.0, 0.3 * T1, T1, 1.7 * T1):
        assert abs(o.funbou(p["fext_func"], 0, t) - t / T1) <= 1e-15 * max(t / T1, 1.0)
    assert abs(o.funbou(0, 0, T1 / 2) - PRESSURE) < 1e-12
    # the top faces' normal is +z: a negative pressure pushes down, on the
    # 4 x 1 mm top minus the four corner shares (0.25 mm^2 / 4 each) that sit
    # on the clamped edge, whose dofs take no load
    ff = o.face_loads(np.zeros_like(ref["coord"]), T1 / 2)
    np.testing.assert_allclose(ff.sum(axis=0), [0.0, 0.0, PRESSURE * (4.0 - 0.25)], atol=1e-12)
    assert (ff[clamped] == 0.0).all()
    assert (ff[np.abs(ref["coord"][:, 2] - LZ) > 1e-9] == 0.0).all()


def test_deck_gravity(tmp_path):
    """GRAVITY: NORM=..., GX=..., GY=..., GZ=... in PROBLEM_DEFINITION gives
    g = NORM times the normalised direction (sld_reaphy.f90:209-223)."""
    from _decks import write_beam_deck
    path, _ = write_beam_deck(str(tmp_path))
    sld = os.path.join(str(tmp_path), "beam.sld.dat")
    text = open(sld).read()
    assert "  PROBLEM_DEFINITION\n" in text
    text = text.replace("  PROBLEM_DEFINITION\n",
                        "  PROBLEM_DEFINITION\n    GRAVITY: NORM=9810.0, GX=0.0, GY=3.0, GZ=-4.0\n")
    open(sld, "w").write(text)
    p = deck.read_case(path)
    np.testing.assert_allclose(p["gravity"], [0.0, 9810.0 * 0.6, -9810.0 * 0.8], rtol=1e-15)
    os.makedirs(str(tmp_path / "plain"))
    assert "gravity" not in deck.read_case(write_beam_deck(str(tmp_path / "plain"))[0])



def _copy_case(tmp_path):
    import shutil
    src = os.path.dirname(CASE)
    dst = str(tmp_path / "case")
    shutil.copytree(src, dst)
    return os.path.join(dst, "sm154_d1C.dat"), dst


def test_deck_node_sets_take_one_node_each(tmp_path):
    """A set id listed for several nodes outputs its highest-numbered node
    (mod_reaset.f90:502-507), and a node listed twice keeps its last set."""
    path, d = _copy_case(tmp_path)
    with open(os.path.join(d, "sm154_d1C.set.dat"), "w") as f:
        f.write("NODE\n  4 1\n  2 1\n  8 2\n  6 2\n  5 2\n  5 1\nEND_NODE\n")
    p = deck.read_case(path)
    assert {k: v.tolist() for k, v in p["node_sets"].items()} == {1: [1, 3, 4], 2: [5, 7]}
    assert p["set_node"] == {1: 4, 2: 7}


def test_deck_rayleigh_damping_and_body_force(tmp_path):
    """RAYLEIGH_DAMPING: RAYALPHA/RAYBETA per material (sld_reaphy.f90:
    1046-1049; 'RAYLE' matches on five letters, absent values are 0) and
    BODYF = function under transient conditions (sld_reabcs.f90:523)."""
    path, d = _copy_case(tmp_path)
    sld = os.path.join(d, "sm154_d1C.sld.dat")
    text = open(sld).read()
    text = text.replace("    DENSITY           = 6.4e-9\n",
                        "    DENSITY           = 6.4e-9\n    RAYLEIGH_DUMPING, RAYALPHA= 100000\n")
    text = text.replace("BOUNDARY_CONDITIONS, TRANSIENT\n",
                        "BOUNDARY_CONDITIONS, TRANSIENT\n  BODYF = 1\n")
    open(sld, "w").write(text)
    p = deck.read_case(path)
    np.testing.assert_array_equal(p["damping"], [[1.0e5, 0.0]])
    assert p["body_func"] == 0
    text = text.replace("RAYLEIGH_DUMPING, RAYALPHA= 100000",
                        "RAYLEIGH_DAMPING: RAYALPHA=2.5, RAYBETA=0.125")
    text = text.replace("BOUNDARY_CONDITIONS, TRANSIENT", "BOUNDARY_CONDITIONS, CONSTANT")
    open(sld, "w").write(text)
    p = deck.read_case(path)
    np.testing.assert_array_equal(p["damping"], [[2.5, 0.125]])
    assert "body_func" not in p           # constant conditions: bodyf_sld stays 0
    q = deck.read_case(CASE)
    assert "damping" not in q and "body_func" not in q
