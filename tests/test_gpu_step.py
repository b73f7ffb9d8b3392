"""Whole explicit step on the GPU (sld_dev_*) against the CPU oracle run in
the same colour order: nodal fields, forces, stresses <= 1e-10 relative in
FP64, history and branch flags exact; determinism; restart; FP32 error."""
import numpy as np
import pytest

from alya_sma_amd import problems
from alya_sma_amd.solid import Solidz
from oracle import oracle_py as O

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def _pair(prob, store=True):
    dev = Solidz(prob)
    if store:
        dev.set_store(True)
    ptr, order = dev.coloring()
    orc = O.StepOracle(prob, elem_order=order, color_ptr=ptr)
    return dev, orc


def _compare(dev, orc, label, quiet=False):
    u, v, a, f = dev.get("u"), dev.get("v"), dev.get("a"), dev.get("fint")
    sdv = dev.get("sdv")
    S = dev.get("stress")
    fl = dev.get("flags")
    # a = (f_ext - f_int)/m is a small difference of large nodal forces: its
    # natural scale is max|f_int| / min(m), not max|a|
    a_scale = max(np.max(np.abs(orc.fint)) / np.min(orc.mass), np.max(np.abs(orc.a)))
    # F formed from coordinates (here ~10 mm over 1 mm elements) carries
    # ~1e-15 absolute rounding, so a stress at strain eps holds ~1e-15/eps of
    # its digits in either code: stresses are compared relative to
    # max(|S|, E 1e-4)
    S_scale = max(np.max(np.abs(orc.S)), 1e-4 * orc.E_max)
    errs = dict(u=_rel(u, orc.u), v=_rel(v, orc.v), a=np.max(np.abs(a - orc.a)) / a_scale,
                fint=_rel(f, orc.fint), S=np.max(np.abs(S - orc.S)) / S_scale)
    if not quiet:
        print(label, {k: f"{x:.1e}" for k, x in errs.items()}, flush=True)
    for k, x in errs.items():
        assert x < TOL, (label, k, x)
    np.testing.assert_array_equal(fl & 3, orc.flags & 3)
    assert _rel(sdv[..., 0], orc.sdv[0, ..., 0]) < 1e-9
    assert _rel(sdv[..., 1:], orc.sdv[0, ..., 1:]) < TOL


def _sync_to_oracle(orc, dev):
    """Copy the device's state into the oracle (both history slots)."""
    orc.u[:] = dev.get("u")
    orc.v[:] = dev.get("v")
    orc.a[:] = dev.get("a")
    sdv = dev.get("sdv")
    orc.sdv[0] = sdv
    orc.sdv[1] = sdv
    orc.s.t = float(dev.get("time")[0])


def _sync_from_oracle(dev, orc):
    """Copy the oracle's state into the device (restart path)."""
    dev.set("u", orc.u)
    dev.set("v", orc.v)
    dev.set("a", orc.a)
    dev.set("sdv", orc.sdv[1])
    dev.set("time", np.array([orc.t]))


@pytest.mark.parametrize("scheme", [0, 1])
def test_bar_10x2x2_full_cycle(scheme):
    """10x2x2 bar, ramp to 6% strain and back, 2000 steps.

    The undamped unloading through the reverse plateau amplifies a 1e-15
    perturbation a hundredfold every ~50 steps near its end (the CPU oracle
    does the same to itself), so the device restarts from the oracle's state
    every 25 steps and the step map is compared over each chunk.  The first
    chunk ends at ~1e-5 strain, where F formed from the 10 mm coordinates keeps
    only ~1e-10 of its digits in either code: comparisons start at step 50
    (strain ~7e-5)."""
    prob = problems.bar_problem()
    r = prob["run"]
    print(f"bar 10x2x2 scheme {scheme}: 2000 steps", flush=True)
    dev, orc = _pair(prob)
    with dev:
        np.testing.assert_allclose(dev.get("mass"), orc.mass, rtol=1e-13)
        for chunk in range(80):
            if chunk:
                _sync_from_oracle(dev, orc)
            dev.step(r["dt"], nsteps=25, scheme=scheme, phi=1.033)
            orc.step(r["dt"], scheme=scheme, phi=1.033, n=25)
            if chunk:
                _compare(dev, orc, f"bar scheme {scheme} step {25 * (chunk + 1)}", quiet=chunk % 20 != 19)
        xs = orc.sdv[0, ..., 0]
        c = dev.counters()
        assert c["kf_AS"] > 0 and c["kf_SA"] > 0
        t = dev.get("time")[0]
        assert abs(t - orc.t) <= 1e-15 * orc.t


def test_bar_10x2x2_free_run_first_1500_steps():
    """Without resynchronisation the trajectories agree to 1e-10 through the
    loading, saturation and start of unloading."""
    prob = problems.bar_problem()
    dt = prob["run"]["dt"]
    dev, orc = _pair(prob)
    with dev:
        dev.step(dt, nsteps=1500)
        orc.step(dt, n=1500)
        _compare(dev, orc, "bar free run 1500")


def test_small_cantilever_with_noise_and_tip_load():
    prob = problems.cantilever_problem(nx=24, ny=4, nz=4, lx=6.0, ly=1.0, lz=1.0, nsteps=600,
                                       traction=700.0)
    r = prob["run"]
    print("cantilever 24x4x4: 600 steps", flush=True)
    dev, orc = _pair(prob)
    with dev:
        dev.step(r["dt"], nsteps=300, scheme=0)
        orc.step(r["dt"], scheme=0, n=300)
        _compare(dev, orc, "cantilever 300")
        dev.step(r["dt"], nsteps=300, scheme=0)
        orc.step(r["dt"], scheme=0, n=300)
        _compare(dev, orc, "cantilever 600")
        assert np.any(orc.flags & 1)


def test_deterministic_bitwise():
    prob = problems.cantilever_problem(nx=24, ny=4, nz=4, lx=6.0, ly=1.0, lz=1.0, nsteps=200,
                                       traction=700.0)
    r = prob["run"]
    outs = []
    for _ in range(2):
        with Solidz(prob) as dev:
            dev.step(r["dt"], nsteps=200)
            outs.append((dev.get("u"), dev.get("fint"), dev.get("sdv")))
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


def test_restart_bitwise():
    prob = problems.bar_problem(nsteps=400)
    dt = prob["run"]["dt"]
    with Solidz(prob) as a:
        a.step(dt, nsteps=400)
        ref = (a.get("u"), a.get("v"), a.get("sdv"))
    with Solidz(prob) as b:
        b.step(dt, nsteps=200)
        state = {k: b.get(k) for k in ("u", "v", "a", "sdv", "time")}
    with Solidz(prob) as c:
        for k, x in state.items():
            c.set(k, x)
        c.step(dt, nsteps=200)
        got = (c.get("u"), c.get("v"), c.get("sdv"))
    for x, y in zip(ref, got):
        assert np.array_equal(x, y)


def test_odd_step_counts_and_eager_path_agree():
    prob = problems.bar_problem(nsteps=200)
    dt = prob["run"]["dt"]
    with Solidz(prob) as a, Solidz(prob) as b:
        a.step(dt, nsteps=37)
        for _ in range(37):
            b.step(dt, nsteps=1)
        assert np.array_equal(a.get("u"), b.get("u"))
        assert np.array_equal(a.get("sdv"), b.get("sdv"))


@pytest.mark.parametrize("assembly", ["gather", "scatter"])
@pytest.mark.parametrize("epb,persistent", [(8, 0), (16, 0), (32, 0), (8, 1), (32, 2)])
def test_tiles_agree(epb, persistent, assembly):
    # 40x8x8 = 2560 elements: persistent launches with 1-2 workgroups per CU
    # still walk several tiles per workgroup on 256 CUs only for epb 8
    prob = problems.cantilever_problem(nx=40, ny=8, nz=8, lx=10.0, ly=2.0, lz=2.0, nsteps=100,
                                       traction=700.0)
    dt = prob["run"]["dt"]
    with Solidz(prob) as a, Solidz(prob) as b:
        b.set_assembly(assembly)
        b.set_tile(epb, xcd_remap=(epb != 16), persistent=persistent)
        a.step(dt, nsteps=100)
        b.step(dt, nsteps=100)
        assert np.array_equal(a.get("u"), b.get("u"))
        assert np.array_equal(a.get("sdv"), b.get("sdv"))


@pytest.mark.parametrize("mode,scheme", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_gather_and_scatter_assembly_bitwise_equal(mode, scheme):
    """The gather sums each node's element corners in colour order, the
    coloured scatter adds them colour by colour: the same additions."""
    prob = problems.cantilever_problem(nx=24, ny=4, nz=4, lx=6.0, ly=1.0, lz=1.0, nsteps=300,
                                       traction=700.0)
    dt = prob["run"]["dt"]
    print(f"assembly gather vs scatter, mode {mode} scheme {scheme}", flush=True)
    with Solidz(prob) as g, Solidz(prob) as s:
        g.set_assembly("gather")
        s.set_assembly("scatter")
        for d in (g, s):
            d.set_mode(mode)
            d.step(dt, nsteps=150, scheme=scheme, phi=1.033)
        for f in ("u", "v", "a", "fint", "sdv", "time"):
            assert np.array_equal(g.get(f), s.get(f)), f
        # switching assembly mid-run keeps the trajectory
        g.set_assembly("scatter")
        s.set_assembly("gather")
        for d in (g, s):
            d.step(dt, nsteps=150, scheme=scheme, phi=1.033)
        for f in ("u", "v", "fint", "sdv"):
            assert np.array_equal(g.get(f), s.get(f)), f
        assert np.any(s.get("sdv")[..., 0] > 0)


def test_element_with_repeated_node_gather_matches_oracle():
    """A hex whose corner 3 sits on corner 2 (a collapsed edge): two lanes
    of one element add to the same node, which only the gather assembly
    orders; the scatter assembly is refused for such meshes."""
    prob = problems.cantilever_problem(nx=8, ny=2, nz=2, lx=4.0, ly=1.0, lz=1.0, nsteps=200,
                                       traction=500.0)
    ln = prob["lnods"].copy()
    ln[5, 3] = ln[5, 2]
    prob["lnods"] = ln
    _, gpvol, _, _ = O.hex8_kinematics(prob["coord"][ln[5]], np.zeros((8, 3)))
    assert np.all(gpvol > 0)
    dt = prob["run"]["dt"]
    from alya_sma_amd._lib import HipError
    dev, orc = _pair(prob)
    with dev:
        with pytest.raises(HipError, match="node twice"):
            dev.set_assembly("scatter")
        dev.step(dt, nsteps=200)
        orc.step(dt, n=200)
        _compare(dev, orc, "repeated node 200")


def _fan_problem(ntheta=12, nr=2, nz=2, radius=1.0, height=1.0):
    """Hexes around an axis (r, theta, z); the innermost ring collapses onto
    one axis node per level, so each inner element lists the axis nodes
    twice and an interior axis node gathers 4 x ntheta element corners."""
    axis = lambda iz: iz
    ring = lambda ir, it, iz: (nz + 1) + (ir - 1) + nr * ((it % ntheta) + ntheta * iz)
    node = lambda ir, it, iz: axis(iz) if ir == 0 else ring(ir, it, iz)
    npoin = (nz + 1) + nr * ntheta * (nz + 1)
    coord = np.zeros((npoin, 3))
    for iz in range(nz + 1):
        coord[axis(iz)] = (0.0, 0.0, height * iz / nz)
        for ir in range(1, nr + 1):
            for it in range(ntheta):
                a = 2 * np.pi * it / ntheta
                r = radius * ir / nr
                coord[ring(ir, it, iz)] = (r * np.cos(a), r * np.sin(a), height * iz / nz)
    lnods = np.array([[node(ir, it, iz), node(ir + 1, it, iz), node(ir + 1, it + 1, iz),
                       node(ir, it + 1, iz), node(ir, it, iz + 1), node(ir + 1, it, iz + 1),
                       node(ir + 1, it + 1, iz + 1), node(ir, it + 1, iz + 1)]
                      for iz in range(nz) for it in range(ntheta) for ir in range(nr)], np.int32)
    fix = np.zeros(npoin, np.uint8)
    fix[np.abs(coord[:, 2]) < 1e-12] = 7
    top = np.abs(coord[:, 2] - height) < 1e-12
    fext0 = np.zeros((npoin, 3))
    fext0[top, 0] = 2.0
    fext0[top, 2] = 3.0
    prob = dict(coord=coord, lnods=lnods, props=problems.NITINOL[None].copy(),
                rho=np.array([problems.RHO_NITINOL]), fixmask=fix,
                func_of_node=np.full(npoin, -1, np.int32), E_of_elem=None, fext0=fext0)
    return prob


def test_fan_with_many_corners_at_a_node_gather_matches_oracle():
    """More than 16 element corners at a node: the gather runs with per-node
    slot ranges instead of padded slots (and with repeated nodes)."""
    prob = _fan_problem()
    assert np.bincount(prob["lnods"].ravel()).max() == 4 * 12
    orc0 = O.StepOracle(dict(prob, func_table=np.zeros((1, 2, 4)), func_npts=np.array([2], np.int32),
                             func_smooth=np.array([1], np.int32), fext_func=-1))
    dt = 0.5 * orc0.critical_dt()
    prob.update(func_table=problems.smooth_table(400 * dt, [[0, 0, 0], [1, 0, 0], [0, 0, 0]])[None],
                func_npts=np.array([3], np.int32), func_smooth=np.array([1], np.int32), fext_func=0)
    for k in range(len(prob["lnods"])):
        _, gpvol, _, _ = O.hex8_kinematics(prob["coord"][prob["lnods"][k]], np.zeros((8, 3)))
        assert np.all(gpvol > 0)
    dev, orc = _pair(prob)
    with dev:
        for chunk in range(4):
            dev.step(dt, nsteps=100)
            orc.step(dt, n=100)
            _compare(dev, orc, f"fan {100 * (chunk + 1)}")
        assert np.abs(orc.u).max() > 0.0


@pytest.mark.parametrize("damped", [False, True])
def test_autotune_leaves_state_unchanged(damped):
    """Every field a step writes -- state, forces, stored stress and flags,
    counters, energy balance, reactions -- reads the same right after the
    autotuner as on an untouched twin, and both go on bitwise equal."""
    prob = problems.cantilever_problem(nx=40, ny=8, nz=8, lx=10.0, ly=2.0, lz=2.0, nsteps=100,
                                       traction=700.0)
    if damped:
        prob["damping"] = np.array([[1.0e5, 0.0]])
    dt = prob["run"]["dt"]
    fields = ("u", "v", "a", "fint", "sdv", "time", "stress", "pk1", "flags", "counters",
              "reactions")
    with Solidz(prob) as a, Solidz(prob) as b:
        for d in (a, b):
            d.set_store(True)
            d.set_energy(True)
            d.step(dt, nsteps=50)
        best = b.autotune(dt, nsteps=2)
        print("autotune:", best, flush=True)
        assert best["elems_per_block"] in (8, 16, 32) and best["ms_per_step"] > 0
        assert best["assembly"] in (("gather",) if damped else ("gather", "scatter"))
        for f in fields:
            assert np.array_equal(a.get(f), b.get(f)), f
        assert a.energy() == b.energy()
        a.step(dt, nsteps=50)
        b.step(dt, nsteps=50)
        for f in fields:
            assert np.array_equal(a.get(f), b.get(f)), f
        assert a.energy() == b.energy()


def test_fp32_mode_error_vs_fp64_oracle():
    prob = problems.bar_problem()
    r = prob["run"]
    dev, orc = _pair(prob)
    with dev:
        dev.set_mode(1)
        dev.step(r["dt"], nsteps=1000, scheme=0)
        orc.step(r["dt"], scheme=0, n=1000)
        S = dev.get("stress")
        eS = np.linalg.norm(S - orc.S) / np.linalg.norm(orc.S)
        eu = np.linalg.norm(dev.get("u") - orc.u) / np.linalg.norm(orc.u)
        exs = np.max(np.abs(dev.get("sdv")[..., 0] - orc.sdv[0, ..., 0]))
        print(f"fp32 mode, bar 1000 steps: stress rel L2 {eS:.2e}, u rel L2 {eu:.2e}, "
              f"xS max abs {exs:.2e}", flush=True)
        assert eS < 1e-3 and eu < 1e-3


def test_batched_uneven_specimens_match_oracle():
    """Batched specimens in small: six bars of uneven size, each with its own material
    row and load function, stepped as one mesh (one element launch) against
    the oracle, restarted from the oracle's state every 25 steps."""
    prob = problems.batched_specimens(n=6, median=80, lo=30, hi=300, nsteps=400)
    dt = prob["run"]["dt"]
    print("batched specimens:", prob["name"], flush=True)
    dev, orc = _pair(prob)
    with dev:
        for chunk in range(16):
            if chunk:
                _sync_from_oracle(dev, orc)
            dev.step(dt, nsteps=25)
            orc.step(dt, n=25)
            _compare(dev, orc, f"batch step {25 * (chunk + 1)}", quiet=chunk % 8 != 7)
        xs = dev.get("sdv")[..., 0].max(axis=1)
        mat = prob["mat_of_elem"]
        assert all(np.any(xs[mat == s] > 0) for s in range(6))   # every specimen transformed


def test_prepare_builds_graphs_and_runs_no_step():
    """sld_dev_prepare builds the captured runs a step call would replay and
    leaves the state as it was; the prepared call then gives bitwise the
    state of an unprepared one (83 steps: runs of 64, 16 and 2, one eager)."""
    prob = problems.bar_problem()
    dt = prob["run"]["dt"]
    print("prepare vs plain step, bar, 83 steps", flush=True)
    with Solidz(prob) as a, Solidz(prob) as b:
        for d in (a, b):
            d.step(dt, nsteps=301)
        before = {f: a.get(f) for f in ("u", "v", "a", "sdv", "time")}
        a.prepare(dt, nsteps=83)
        a.prepare(dt, nsteps=0)
        for f, v in before.items():
            np.testing.assert_array_equal(a.get(f), v)
        a.step(dt, nsteps=83)
        b.step(dt, nsteps=83)
        for f in ("u", "v", "a", "sdv", "time"):
            np.testing.assert_array_equal(a.get(f), b.get(f))


def test_batched_specimens_surrogate_mode_one_step():
    """Surrogate mode on batched specimens, each with its own material row
    (the network's features carry each element's E and material): one
    surrogate step from an FP64 state in the loading phase stays within
    1e-3 (relative L2 stress) of the oracle's FP64 step, with points listed
    for the network and every accepted point's history within the
    acceptance tolerance of the oracle's."""
    from alya_sma_amd import surrogate as S
    net = S.load_default()
    assert net is not None, "the committed network is missing"
    prob = problems.batched_specimens(n=6, median=80, lo=30, hi=300, nsteps=400)
    dt = prob["run"]["dt"]
    print("batched specimens: one surrogate step after 150 FP64 steps", flush=True)
    dev, orc = _pair(prob)
    with dev:
        dev.step(dt, nsteps=150)
        _sync_to_oracle(orc, dev)
        dev.set_surrogate(net, 1e-4)
        dev.set_mode(2)
        c0 = dev.counters()
        dev.step(dt, nsteps=1)
        orc.step(dt, n=1)
        c1 = dev.counters()
        Sd = dev.get("stress")
        l2 = np.linalg.norm(Sd - orc.S) / np.linalg.norm(orc.S)
        used = c1["surrogate_used"] - c0["surrogate_used"]
        seen = used + c1["surrogate_fallback"] - c0["surrogate_fallback"]
        fl = dev.get("flags")
        acc = ((fl & 0x20) > 0) & ((fl & 3) == (orc.flags & 3))
        dxs = np.abs(dev.get("sdv")[..., 0] - orc.sdv[0, ..., 0])
        print(f"  stress rel L2 {l2:.2e}, transforming {int(np.count_nonzero(orc.flags & 3))}, "
              f"seen {seen}, accepted {used}", flush=True)
        assert np.count_nonzero(orc.flags & 3) > 0 and seen > 0
        assert l2 < 1e-3
        if acc.any():
            assert dxs[acc].max() <= 1e-4 + 1e-6


def test_batched_specimens_persistent_launch_bitwise():
    prob = problems.batched_specimens(n=12, median=200, lo=50, hi=900, nsteps=200)
    dt = prob["run"]["dt"]
    with Solidz(prob) as a, Solidz(prob) as b:
        b.set_tile(8, xcd_remap=True, persistent=2)
        a.step(dt, nsteps=120)
        b.step(dt, nsteps=120)
        for f in ("u", "v", "sdv"):
            assert np.array_equal(a.get(f), b.get(f)), f


def test_tube_in_tissue_two_materials_match_oracle():
    """Tube in tissue in small: SMA tube (model 160) in neo-Hookean tissue (model
    102), inner surface crimped radially, against the oracle (restarted from
    its state every 25 steps), plus the critical step of the two materials."""
    prob = problems.cylinder_problem(ntheta=16, nz=6, nr_tube=2, nr_tissue=3, nsteps=400)
    dt = prob["run"]["dt"]
    print("tube in tissue:", prob["name"], flush=True)
    dev, orc = _pair(prob)
    with dev:
        assert abs(dev.critical_dt() - orc.critical_dt()) <= 1e-13 * orc.critical_dt()
        # the elements are stored grouped by law: within a colour the SMA
        # elements come before the tissue
        ptr, order = dev.coloring()
        laws = prob["mat_of_elem"][order]
        for c in range(len(ptr) - 1):
            assert np.all(np.diff(laws[ptr[c]:ptr[c + 1]]) >= 0)
        for chunk in range(16):
            if chunk:
                _sync_from_oracle(dev, orc)
            dev.step(dt, nsteps=25)
            orc.step(dt, n=25)
            _compare(dev, orc, f"tube step {25 * (chunk + 1)}", quiet=chunk % 8 != 7)
        xs = dev.get("sdv")[..., 0]
        mat = prob["mat_of_elem"]
        assert xs[mat == 0].max() > 0          # the tube transforms under the crimp
        assert np.all(xs[mat == 1] == 0)       # the tissue has no history


def test_tube_in_tissue_scatter_runs_each_elements_law():
    """The coloured scatter's colours cut across the law groups, so its
    launches run the kernel that takes each element's own law; the gather
    runs the SMA-only and the neo-Hookean-only kernels over the two ranges.
    Same arithmetic per point, same additions per node: bitwise equal, in
    FP64 and FP32, and the scatter matches the oracle."""
    prob = problems.cylinder_problem(ntheta=16, nz=6, nr_tube=2, nr_tissue=3, nsteps=400)
    dt = prob["run"]["dt"]
    print("tube in tissue: gather vs scatter", flush=True)
    for mode in (0, 1):
        with Solidz(prob) as g, Solidz(prob) as sc:
            sc.set_assembly("scatter")
            for d in (g, sc):
                d.set_mode(mode)
                d.step(dt, nsteps=350)
            for f in ("u", "v", "a", "fint", "sdv"):
                assert np.array_equal(g.get(f), sc.get(f)), (mode, f)
            assert sc.get("sdv")[..., 0].max() > 0
    dev, orc = _pair(prob)
    with dev:
        dev.set_assembly("scatter")
        for chunk in range(8):
            if chunk:
                _sync_from_oracle(dev, orc)
            dev.step(dt, nsteps=25)
            orc.step(dt, n=25)
        _compare(dev, orc, "tube scatter 200")


def test_tube_in_tissue_surrogate_mode_one_step():
    """Surrogate mode on a mixed (law-grouped) mesh: the prepass covers the
    SMA range, the step runs the SMA range in the surrogate kernel (the
    network's update where the check accepts it, the exact map elsewhere)
    and the tissue in the neo-Hookean-only kernel of the FP32 mode, so the
    tissue's stresses equal an FP32-mode step's bitwise.  One step from an
    FP64 state in the crimp stays within 1e-3 (relative L2 stress) of the
    oracle's FP64 step, and the tissue keeps no history."""
    from alya_sma_amd import surrogate as S
    net = S.load_default()
    assert net is not None, "the committed network is missing"
    prob = problems.cylinder_problem(ntheta=16, nz=6, nr_tube=2, nr_tissue=3, nsteps=400)
    dt = prob["run"]["dt"]
    print("tube in tissue: one surrogate step after 350 FP64 steps", flush=True)
    dev, orc = _pair(prob)
    with dev, Solidz(prob) as twin:
        dev.step(dt, nsteps=350)
        _sync_to_oracle(orc, dev)
        twin.set_store(True)
        for f in ("u", "v", "a", "sdv", "time"):
            twin.set(f, dev.get(f))
        twin.set_mode(1)
        twin.step(dt, nsteps=1)
        dev.set_surrogate(net, 1e-4)
        dev.set_mode(2)
        c0 = dev.counters()
        dev.step(dt, nsteps=1)
        orc.step(dt, n=1)
        c1 = dev.counters()
        Sd = dev.get("stress")
        l2 = np.linalg.norm(Sd - orc.S) / np.linalg.norm(orc.S)
        listed = (c1["surrogate_used"] - c0["surrogate_used"]) + \
            (c1["surrogate_fallback"] - c0["surrogate_fallback"])
        print(f"  stress rel L2 {l2:.2e}, transforming {int(np.count_nonzero(orc.flags & 3))}, "
              f"listed {listed}", flush=True)
        assert np.count_nonzero(orc.flags & 3) > 0 and listed > 0
        assert l2 < 1e-3
        mat = prob["mat_of_elem"]
        assert np.all(dev.get("sdv")[..., 0][mat == 1] == 0)
        np.testing.assert_array_equal(Sd[mat == 1], twin.get("stress")[mat == 1])


def test_neo_hookean_only_mesh_matches_oracle():
    """Every element neo-Hookean (law 102): the whole mesh is the
    neo-Hookean-only launch.  A soft cantilever with a small tip load."""
    prob = problems.cantilever_problem(nx=16, ny=4, nz=4, lx=4.0, ly=1.0, lz=1.0, nsteps=400,
                                       traction=0.002, E_noise=0.0)
    prob["props"] = problems.TISSUE[None].copy()
    prob["law_of_mat"] = np.array([problems.LAW_NEOHOOKEAN], np.int32)
    prob["rho"] = np.array([problems.RHO_TISSUE])
    prob["E_of_elem"] = None
    with Solidz(prob) as probe:
        dt = 0.5 * probe.critical_dt()
    # the load ramps 0 -> 1 -> 0 over the 400 steps of this step size
    prob["func_table"] = problems.smooth_table(400 * dt, [[0, 0, 0], [1, 0, 0], [0, 0, 0]])[None]
    dev, orc = _pair(prob)
    with dev:
        assert abs(dev.critical_dt() - orc.critical_dt()) <= 1e-13 * orc.critical_dt()
        for chunk in range(4):
            dev.step(dt, nsteps=100, scheme=0)
            orc.step(dt, scheme=0, n=100)
            _compare(dev, orc, f"neo-Hookean cantilever {100 * (chunk + 1)}")
        assert np.abs(orc.u).max() > 1e-3
        assert np.all(dev.get("sdv") == 0.0)


def test_tube_in_tissue_fp32_one_step_error():
    prob = problems.cylinder_problem(ntheta=24, nz=8, nr_tube=2, nr_tissue=4, nsteps=400)
    dt = prob["run"]["dt"]
    with Solidz(prob) as ref, Solidz(prob) as f32:
        ref.set_store(True)
        f32.set_mode(1)
        f32.set_store(True)
        ref.step(dt, nsteps=199)
        for k in ("u", "v", "a", "sdv", "time"):
            f32.set(k, ref.get(k))
        ref.step(dt, nsteps=1)
        f32.step(dt, nsteps=1)
        S, Sr = f32.get("stress"), ref.get("stress")
        err = np.linalg.norm(S - Sr) / np.linalg.norm(Sr)
        print(f"tube in tissue, fp32 one-step stress error {err:.2e}", flush=True)
        assert err < 1e-3


@pytest.mark.parametrize("scheme", [0, 1])
@pytest.mark.parametrize("assembly", ["gather", "scatter"])
@pytest.mark.parametrize("mode", [0, 1])
def test_fused_graph_runs_equal_eager_steps(scheme, assembly, mode):
    """A run of 64 + 2 x 16 + 2 + 1 steps replays captured graphs whose
    correctors also predict the next step; one step per call runs every
    kernel on its own.  State, energies and flag counters are bitwise equal."""
    prob = problems.cantilever_problem(nx=40, ny=4, nz=4, lx=5.0, ly=0.5, lz=0.5, traction=800.0,
                                       nsteps=400)
    r = prob["run"]
    out = []
    for per_call in (99, 1):
        with Solidz(prob) as dev:
            dev.set_assembly(assembly)
            dev.set_mode(mode)
            dev.set_store(True)
            dev.set_energy(True)
            for _ in range(99 // per_call):
                dev.step(r["dt"], nsteps=per_call, scheme=scheme, phi=r["phi"])
            out.append({k: dev.get(k) for k in ("u", "v", "a", "fint", "sdv", "stress", "flags", "time")}
                       | {"energy": np.array(list(dev.energy().values())),
                          "counters": np.array(list(dev.counters().values()))})
    print(f"fused vs eager scheme {scheme} {assembly} mode {mode}: "
          f"max |u| {np.abs(out[1]['u']).max():.3e}", flush=True)
    for k in out[0]:
        np.testing.assert_array_equal(out[0][k], out[1][k], err_msg=k)


def test_pk1_and_reactions_fields():
    """First Piola P = F S per Gauss point (gppio_sld) against the oracle's
    F (restated kinematics) times its S, and P F^T symmetric (= tau); the
    reactions field is -f_int on prescribed dofs and zero elsewhere
    (frxid_sld), balancing the prescribed end of the bar."""
    prob = problems.bar_problem()
    r = prob["run"]
    dev, orc = _pair(prob)
    with dev:
        dev.step(r["dt"], nsteps=600, scheme=0, phi=1.033)
        orc.step(r["dt"], scheme=0, phi=1.033, n=600)
        P = dev.get("pk1")
        S = dev.get("stress").reshape(-1, 8, 3, 3).transpose(0, 1, 3, 2)   # col-major -> [i][j]
        Pm = P.reshape(-1, 8, 3, 3).transpose(0, 1, 3, 2)
        u = dev.get("u")
        X = prob["coord"]
        err = 0.0
        for e, nodes in enumerate(prob["lnods"]):
            _, _, F, _ = O.hex8_kinematics(X[nodes], u[nodes])
            ref = F @ S[e]
            err = max(err, np.abs(Pm[e] - ref).max())
            tau = Pm[e] @ F.transpose(0, 2, 1)
            assert np.abs(tau - tau.transpose(0, 2, 1)).max() <= 1e-9 * np.abs(tau).max()
        assert err <= 1e-10 * np.abs(Pm).max(), err
        R = dev.get("reactions")
        fint = dev.get("fint")
        fixed = np.stack([(prob["fixmask"] >> k) & 1 for k in range(3)], axis=1).astype(bool)
        np.testing.assert_array_equal(R[~fixed], 0.0)
        np.testing.assert_array_equal(R[fixed], -fint[fixed])
        # x-reactions of the two x-faces (the clamped one and the prescribed one)
        # balance up to the inertia of the bar
        x = X[:, 0]
        lo, hi = R[np.abs(x) < 1e-9, 0].sum(), R[np.abs(x - x.max()) < 1e-9, 0].sum()
        assert lo * hi < 0.0


@pytest.mark.parametrize("eps", [1e-120, 1e-90, 1e-20])
def test_fp64_step_stress_of_tiny_strains(eps):
    """A uniform uniaxial strain eps on a free bar, one FP64 step with zero
    velocity and acceleration (the predictor leaves u as set): every Gauss
    point carries the linear-elastic stress S_xx = (lambda + 2 mu) eps,
    S_yy = S_zz = lambda eps to 1e-10 (the law's nonlinear terms are of
    relative order eps).  At 1e-120 and 1e-90 the reference's
    law returns NaN (its eigen-solver underflows, tests/test_oracle_law.py);
    the device's FP64 element step forms the Hencky strain by its series
    and stays defined there."""
    prob = problems.bar_problem(nsteps=100)
    prob["fixmask"] = np.zeros_like(prob["fixmask"])
    prob["func_of_node"] = np.full_like(prob["func_of_node"], -1)
    E, nu = prob["props"][0, 0], prob["props"][0, 1]
    K, G = E / (3.0 * (1.0 - 2.0 * nu)), E / (2.0 * (1.0 + nu))
    lam = K - 2.0 * G / 3.0
    with Solidz(prob) as dev:
        dev.set_store(True)
        u = np.zeros((dev.npoin, 3))
        u[:, 0] = eps * prob["coord"][:, 0]
        dev.set("u", u)
        dev.step(prob["run"]["dt"], nsteps=1)
        assert np.array_equal(dev.get("u"), u)
        S = dev.get("stress").reshape(-1, 9)
        fl = dev.get("flags")
    sxx = (lam + 2.0 * G) * eps
    print(f"eps {eps:.0e}: S_xx {S[:, 0].min():.6e}..{S[:, 0].max():.6e} expected {sxx:.6e}",
          flush=True)
    assert np.isfinite(S).all()
    assert np.max(np.abs(S[:, 0] - sxx)) < 1e-10 * sxx
    for k in (4, 8):
        assert np.max(np.abs(S[:, k] - lam * eps)) < 1e-10 * sxx
    for k in (1, 2, 3, 5, 6, 7):
        assert np.max(np.abs(S[:, k])) < 1e-10 * sxx
    assert not np.any(fl & 3)
