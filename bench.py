"""Benchmark: Gauss-point updates per second of the explicit SMA step.

Workload (N = 1): the SMA cantilever, hex8 400x20x20 = 160 000 elements
(1.28 M Gauss points) over 50 x 2.5 x 2.5 mm, clamped at x = 0, tip traction
ramped up and released over a 5000-step central-difference load cycle,
per-element Young's modulus with seeded +-2% noise.  Inputs are resident in
HBM before the timed region; one step = predictor, element kernel(s),
corrector.

    python bench.py [--steps K] [--warmup W] [--skip S] [--gpus 1]
                    [--workload cantilever|specimens|cylinder]

Timed window.  Each precision mode first runs S steps untimed (default: 40%
of the load cycle, 2000 steps on the cantilever, well inside the phase in
which Gauss points transform), then the autotuner picks the launch shape
(state restored), then W warm-up steps, then K steps timed as one block
between two synchronisations (default K: the rest of the load cycle).  The
window, its share of transforming point updates and the surrogate's
accepted updates are reported.

`value` is the FP64 mode, the reference's precision.  The FP32 and
surrogate modes are reported under `modes` with their throughput, their
ratio to FP64 and their stress error: the relative L2 (and the largest)
difference of the Gauss-point PK2 stress after one step taken from the FP64
trajectory's state, at ten checkpoints in the window (max over
checkpoints), the number of points whose branch flags differ from the FP64
step's, and the difference between whole trajectories at the window's end
beside the same difference between two FP64 runs whose states differ by
1e-13 (relative, in u at step 20) -- the sensitivity of the undamped cycle.

Roofline of the element kernel, per mode: after the window, four eager steps
with HIP events on the handle's stream (the element launch repeated 8x
back to back inside its bracket) give the launch duration; bytes are the
algorithmic bytes of one launch (per Gauss point 48 B of history read and
written, 4 B of connectivity, 1 B of per-element modulus; per node 48 B of
coordinates and displacement read and the internal force written), FLOPs
the frozen per-Gauss-point counts of the instrumented CPU oracle
(oracle/flops_per_gp.txt) times the window's point mix.  `traffic` is the
measured HBM bytes per launch from the committed rocprofv3 PMC summary.

`cpu_baseline`: the FP64 C oracle from the state at the start of the
window, on all of the host cores this process may use and on one thread.
"""
import argparse
import json
import math
import os
import re
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gauss-pt updates/s per explicit step; surrogate vs plain at ≤1e-3 stress err"
PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "sld_element_hbm_traffic.txt")
MODES = {"fp64": 0, "fp32": 1, "surrogate": 2}

# Algorithmic FLOPs (add + mul + div; sqrt, log and pow apart) of the
# reference algorithm, counted once by instrumenting the CPU oracle
# (`make -C oracle flopcount && oracle/_build/flopcount`, output frozen in
# oracle/flops_per_gp.txt; tests/test_flopcount.py checks these against it).
FLOPS_PER_GP = {
    "kinematics": 544.0,    # Jacobian, inverse, gpcar, F, det F
    "law_elastic": 612.0,   # sm160 without a branch (6.2 sqrt, 4 log, 1 pow, 17.5 div)
    "law_AS": 695.0,        # austenite -> martensite, one Newton update
    "law_SA": 701.0,        # martensite -> austenite, one Newton update
    "law_nh102": 136.0,     # neo-Hookean law 102
    "force": 354.0,         # P = F S, element internal force, assembly
}
FLOPS_PER_NODE = 34.5       # predictor, solve, corrector


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: the rest of the load cycle after skip + warmup)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--skip", type=int, default=None,
                    help="untimed steps before the warm-up (default: 40%% of the load cycle)")
    ap.add_argument("--workload", choices=["cantilever", "specimens", "cylinder"], default="cantilever")
    ap.add_argument("--modes", default="fp64,fp32,surrogate")
    ap.add_argument("--sur-tol", type=float, default=1e-4,
                    help="surrogate acceptance: distance in xS to the return map's solution")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="CPU baseline sample length per thread count")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the error pass")
    ap.add_argument("--tune", choices=["auto", "none"], default="auto",
                    help="auto: run the library autotuner after the untimed steps")
    ap.add_argument("--assembly", choices=["gather", "scatter"], default=None,
                    help="fix the assembly (disables the autotuner)")
    ap.add_argument("--epb", type=int, default=0, help="elements per block (disables the autotuner)")
    ap.add_argument("--copy-probe-mb", type=int, default=256,
                    help="size of the calibration copy kernel run (HBM counter reference)")
    return ap.parse_args()


def kernel_name(mode, cfg, mixed=False):
    """The element kernel that carries the SMA elements.  Its law range: 1 =
    SMA only (every launch over a mesh without neo-Hookean elements, and the
    SMA range of a law-grouped mesh in the gather assembly, which a
    neo-Hookean-only launch, range 2, follows, in every mode); 0 = each
    element's own law (a mixed mesh in the scatter assembly).  No Rayleigh
    damping."""
    r = "double" if mode == "fp64" else "float"
    sur = 2 if mode == "surrogate" else 0
    ga = "true" if cfg["assembly"] == "gather" else "false"
    laws = 0 if mixed and cfg["assembly"] != "gather" else 1
    return f"sld_element<{r}, {cfg['elems_per_block']}, false, {sur}, {ga}, {laws}, false>"


def _sma_mask(prob):
    from alya_sma_amd.problems import LAW_SMA
    ne = prob["lnods"].shape[0]
    if prob.get("law_of_mat") is None:
        return np.ones(ne, bool)
    mat = prob.get("mat_of_elem")
    law = np.asarray(prob["law_of_mat"])[mat if mat is not None else np.zeros(ne, np.int64)]
    return law == LAW_SMA


def algorithmic_bytes(prob, mode):
    """Bytes one element phase over all elements must move at least: per
    Gauss point the history read and written (3 + 3 doubles, 48 B; SMA
    elements only), connectivity 4 B and, where the workload has one, the
    per-element Young's modulus (8 B per element, 1 B per point); per node
    the coordinates and displacement read (48 B) and the internal force
    written (3 x 8 B in FP64, 3 x 4 B in FP32)."""
    fb = 8 if mode == "fp64" else 4
    sma = _sma_mask(prob)
    per_el = 8 * (4.0 + np.where(sma, 48.0, 0.0))
    if prob.get("E_of_elem") is not None:
        per_el = per_el + 8.0
    npn = prob["coord"].shape[0]
    return float(per_el.sum() + npn * (48 + 3 * fb))


def design_bytes(prob, mode, assembly):
    """Bytes the launch moves by this design's layout (for comparison with
    the algorithmic figure): per Gauss point 48 B of history (SMA), 4 B of
    connectivity, 2 B of the element's bulk and shear moduli, and with the
    gather assembly the corner's slot index 4 B and its force written
    (3 x 8 B or 3 x 4 B) in place of the nodal force; per node 48 B."""
    fb = 8 if mode == "fp64" else 4
    sma = _sma_mask(prob)
    per_el = 8 * (6.0 + np.where(sma, 48.0, 0.0))
    npn = prob["coord"].shape[0]
    if assembly == "gather":
        return float(per_el.sum() + per_el.size * 8 * (4 + 3 * fb) + npn * 48)
    return float(per_el.sum() + npn * (48 + 6 * fb))


def algorithmic_flops(prob, n_as, n_sa):
    """FLOPs of one step's element phase: every point's kinematics and
    force, the law by its branch (n_as / n_sa points per step on the
    loading / unloading branch), the neo-Hookean law on law-102 elements."""
    f = FLOPS_PER_GP
    sma = _sma_mask(prob)
    n_sma, n_nh = 8 * int(sma.sum()), 8 * int((~sma).sum())
    base = (n_sma + n_nh) * (f["kinematics"] + f["force"])
    law = n_sma * f["law_elastic"] + n_nh * f["law_nh102"]
    law += n_as * (f["law_AS"] - f["law_elastic"]) + n_sa * (f["law_SA"] - f["law_elastic"])
    return float(base + law)


def measured_traffic(workload, kernel):
    """HBM bytes per element phase of `kernel` on `workload` from the committed
    rocprofv3 PMC summary (FETCH_SIZE and WRITE_SIZE scaled by a plain copy
    kernel): on a law-grouped mesh the phase is two launches, `kernel` over
    the SMA range (law range 1) and its neo-Hookean-only variant (2), whose
    bytes add.  None when the pair was not measured."""
    nh = kernel.replace(", 1, false>", ", 2, false>")
    found = {}
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                m = re.match(r"traffic_bytes_per_launch\s+(\S+)\s+(.+?)\s+=\s+([0-9.eE+]+)\s*$",
                             line.strip())
                if m and m.group(1) == workload and m.group(2) in (kernel, nh):
                    found[m.group(2)] = float(m.group(3))
    except OSError:
        pass
    return sum(found.values()) if kernel in found else None


def host_cores():
    """(cores this process may run on -- what `nproc` prints -- and the
    cgroup CPU quota if one is set)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()
        if q != "max":
            quota = int(q) / int(period)
    except (OSError, ValueError):
        pass
    return n, quota


def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        return [l.split(":", 1)[1].strip() for l in out.splitlines() if l.startswith("Model name")][0]
    except Exception:
        return "unknown"


def cpu_baseline(prob, state, budget_s):
    """The FP64 C oracle (OpenMP over the elements of each colour) from the
    device state at the start of the timed window: on every core the
    process may use (capped by a cgroup quota), then on one thread."""
    from oracle import oracle_py as O
    from alya_sma_amd import _lib
    lnods = prob["lnods"]
    color = np.zeros(lnods.shape[0], dtype=np.int32)
    nc = _lib.lib().sld_color_elements(lnods.shape[0], prob["coord"].shape[0],
                                       lnods.ctypes.data, color.ctypes.data)
    order = np.argsort(color, kind="stable").astype(np.int32)
    ptr = np.concatenate([[0], np.cumsum(np.bincount(color, minlength=nc))]).astype(np.int32)
    orc = O.StepOracle(prob, elem_order=order, color_ptr=ptr, keep_stress=False)
    dt = prob["run"]["dt"]
    ngp = lnods.shape[0] * 8
    nproc, quota = host_cores()
    threads = nproc if quota is None else max(1, min(nproc, int(math.ceil(quota))))
    res = {}
    # a generic state: the device's FP64 step keeps strains of 1e-162..1e-77
    # (ahead of a loading wave, zero shears of symmetric specimens) where the
    # reference's law -- and so the oracle -- returns NaN and stops; a
    # perturbation of 1e-15 of the largest displacement lifts them out of
    # that band without changing the work of a step
    rng = np.random.default_rng(1)
    u0 = state["u"] + 1e-15 * np.abs(state["u"]).max() * rng.standard_normal(state["u"].shape)
    for nt in (threads, 1):
        O.set_threads(nt)
        orc.u[:], orc.v[:], orc.a[:] = u0, state["v"], state["a"]
        orc.sdv[0], orc.sdv[1] = state["sdv"], state["sdv"]
        orc.s.t = float(state["time"][0])
        if orc.step(dt, scheme=0):                # first step outside the sample (page-in)
            raise RuntimeError(f"CPU baseline: the oracle stopped its run {orc.halt}")
        n, t0 = 0, time.perf_counter()
        while True:
            if orc.step(dt, scheme=0):
                # a stopped run returns at once: the sample would not be a step
                raise RuntimeError(f"CPU baseline: the oracle stopped its run {orc.halt}")
            n += 1
            el = time.perf_counter() - t0
            if el >= budget_s or n >= 400:
                break
        res[nt] = (ngp * n / el, n, el)
    cpu = cpu_model()
    v, n, el = res[threads]
    v1, n1, el1 = res[1]
    where = f"from the state at the start of the timed window, FP64 C oracle (oracle/sld_step.c, " \
            f"gcc -O3, OpenMP over the elements of each colour)"
    return {"value": v, "unit": "GP-updates/s", "cores": threads, "kind": "port",
            "sample": f"{n} steps {where}, {el:.1f} s wall on {threads} threads of {cpu}",
            "single_thread": {"value": v1, "cores": 1,
                              "sample": f"{n1} steps, {el1:.1f} s wall on 1 thread"},
            "cpu_model": cpu, "nproc": nproc, "cgroup_cpu_quota": quota,
            "calibration": "the oracle's law is 1.18x faster than the reference's compiled "
                           "eval_sld_stress_model_160 on one core (oracle/calibrate_law_speed.py)"}


def surrogate_network():
    """The committed trained network, or one trained here (not timed) when
    the file is absent."""
    from alya_sma_amd import surrogate as S
    s = S.load_default()
    if s is None:
        print("# training the surrogate (no committed weights)", file=sys.stderr, flush=True)
        s = S.train_default(n_samples=4_000_000, steps=6000,
                            log=lambda m: print("# " + m, file=sys.stderr, flush=True))
    return s


def make_device(prob, mode, sur=None, tol=1e-4):
    from alya_sma_amd.solid import Solidz
    d = Solidz(prob)
    if mode == "surrogate":
        d.set_surrogate(sur, tol)
    d.set_mode(MODES[mode])
    return d


def accuracy(prob, modes, start, K, cfgs, sur=None, tol=1e-4):
    """One-step stress error at ten checkpoints of the window [start,
    start + K), trajectory error at its end, FP64 self-sensitivity (see the
    module docstring)."""
    dt = prob["run"]["dt"]

    def make(mode, store):
        d = make_device(prob, mode, sur, tol)
        c = cfgs[mode]
        d.set_assembly(c["assembly"])
        d.set_tile(c["elems_per_block"], c["xcd_remap"], c["persistent"])
        if store:
            d.set_store(True)
        return d

    rel = lambda a, b: float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
    relmax = lambda a, b: float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
    others = [m for m in modes if m != "fp64"]
    ref = make("fp64", True)
    pert = make("fp64", True)
    traj = {m: make(m, True) for m in others}
    probe = {m: make(m, True) for m in others}
    first = min(20, start)
    for d in [ref, pert] + list(traj.values()):
        d.step(dt, nsteps=first)
    pert.set("u", pert.get("u") * (1.0 + 1e-13))
    for d in [ref, pert] + list(traj.values()):
        d.step(dt, nsteps=start - first)
    one, onemax = {m: 0.0 for m in others}, {m: 0.0 for m in others}
    # points whose branch flags differ from the FP64 step's, and the largest
    # difference over the points that took the same branch
    flips, same_max = {m: 0 for m in others}, {m: 0.0 for m in others}
    done = 0
    chunk = max(1, K // 10)
    while done < K:
        n = min(chunk, K - done)
        ref.step(dt, nsteps=n - 1)
        if others:
            st = {f: ref.get(f) for f in ("u", "v", "a", "sdv", "time")}
        ref.step(dt, nsteps=1)
        S = ref.get("stress")
        fl = (ref.get("flags") & 3).reshape(-1)
        for m in others:
            for f, v in st.items():
                probe[m].set(f, v)
            probe[m].step(dt, nsteps=1)
            Sm = probe[m].get("stress")
            one[m] = max(one[m], rel(Sm, S))
            onemax[m] = max(onemax[m], relmax(Sm, S))
            same = (probe[m].get("flags") & 3).reshape(-1) == fl
            flips[m] += int(np.count_nonzero(~same))
            d9 = np.abs(Sm - S).reshape(-1, 9)[same]
            if d9.size:
                same_max[m] = max(same_max[m], float(d9.max()) / max(float(np.abs(S).max()), 1e-300))
        for d in [pert] + list(traj.values()):
            d.step(dt, nsteps=n)
        done += n
        print(f"# accuracy pass: {done} of {K} window steps", file=sys.stderr, flush=True)
    S = ref.get("stress")
    out = {"fp64": {"stress_rel_l2_one_step": 0.0, "stress_rel_max_one_step": 0.0,
                    "stress_rel_l2_trajectory_end": 0.0}}
    for m in others:
        out[m] = {"stress_rel_l2_one_step": one[m], "stress_rel_max_one_step": onemax[m],
                  "stress_rel_max_one_step_same_branch": same_max[m],
                  "branch_flips_one_step": flips[m],
                  "stress_rel_l2_trajectory_end": rel(traj[m].get("stress"), S)}
    sens = rel(pert.get("stress"), S)
    for d in [ref, pert] + list(traj.values()) + list(probe.values()):
        d.close()
    return out, sens


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}): the step is "
              "measured on one GPU's share of the mesh; a multi-GPU split of the "
              "elements is not part of this library.", flush=True)
        sys.exit(2)
    from alya_sma_amd import problems
    from alya_sma_amd.device import device_count, synchronize

    if device_count() < 1:
        sys.exit("bench.py needs a HIP device")
    if args.workload == "cantilever":
        prob = problems.cantilever_problem()
        workload = ("SMA cantilever hex8 400x20x20 (160000 elements, 1280000 Gauss points), "
                    "50x2.5x2.5 mm, tip traction load-unload, E +-2% noise, central difference")
    elif args.workload == "cylinder":
        prob = problems.cylinder_problem()
        workload = prob["name"] + ", radial crimp and release of the inner surface, central difference"
    else:
        prob = problems.batched_specimens()
        workload = (prob["name"] + ", lognormal sizes (median 8000, sigma 0.7, 2k-60k), aspect "
                    "5-40, props +-10%, displacement amplitude 3-7%, central difference")
    dt = prob["run"]["dt"]
    cycle = prob["run"]["nsteps"]
    nelem = prob["lnods"].shape[0]
    ngp = nelem * 8
    W = args.warmup
    skip = args.skip if args.skip is not None else int(0.4 * cycle)
    K = args.steps if args.steps is not None else max(1, cycle - skip - W)
    start = skip + W
    modes = [m for m in args.modes.split(",") if m]
    if "fp64" not in modes:
        modes = ["fp64"] + modes
    tune = args.tune == "auto" and not args.assembly and not args.epb
    results, cfgs, state0 = {}, {}, None
    sur = surrogate_network() if "surrogate" in modes else None
    for mode in modes:
        dev = make_device(prob, mode, sur, args.sur_tol)
        if args.assembly:
            dev.set_assembly(args.assembly)
        if args.epb:
            dev.set_tile(args.epb)
        dev.step(dt, nsteps=skip)
        cfg = {"assembly": args.assembly or "gather", "elems_per_block": args.epb or 8,
               "xcd_remap": 1, "persistent": 0}
        if tune:
            cfg = dev.autotune(dt, nsteps=3)
            cfg.pop("ms_per_step")
        cfgs[mode] = cfg
        dev.step(dt, nsteps=W)
        # the timed call's captured graphs are built here, untimed (a caller
        # stepping on keeps them); the timed block replays them
        dev.prepare(dt, nsteps=K)
        dev.sync()
        if mode == "fp64" and not args.no_cpu:
            state0 = {f: dev.get(f) for f in ("u", "v", "a", "sdv", "time")}
        c0 = dev.counters()
        # the K timed steps: one block between two synchronisations
        synchronize()
        t0 = time.perf_counter()
        dev.step(dt, nsteps=K)
        dev.sync()
        synchronize()
        el = time.perf_counter() - t0
        c1 = dev.counters()
        # per-kernel HIP-event times on the state at the window's end
        tk = dev.time_kernels(dt, nsteps=4)
        ms = 1e3 * el / K
        n_as = (c1["kf_AS"] - c0["kf_AS"]) / K
        n_sa = (c1["kf_SA"] - c0["kf_SA"]) / K
        kname = kernel_name(mode, cfg, mixed=not _sma_mask(prob).all())
        launch_s = tk["element_ms"] * 1e-3
        nbytes = algorithmic_bytes(prob, mode)
        achieved = nbytes / launch_s / 1e9
        flops = algorithmic_flops(prob, n_as, n_sa)
        tflops = flops / launch_s / 1e12
        res = {"value": ngp / (ms * 1e-3), "ms_per_step": ms, "launch": cfg,
               "kernel_ms": {k: round(v, 5) for k, v in tk.items()
                             if mode == "surrogate" or not k.startswith("sur_")},
               "transforming_gp_updates": int(c1["kf_AS"] + c1["kf_SA"] - c0["kf_AS"] - c0["kf_SA"]),
               "transforming_fraction": (n_as + n_sa) / ngp,
               "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS,
                            "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS,
                            "traffic": measured_traffic(args.workload, kname), "kernel": kname,
                            "algorithmic_bytes_per_launch": nbytes,
                            "design_bytes_per_launch": design_bytes(prob, mode, cfg["assembly"]),
                            "launch_ms": tk["element_ms"]},
               "roofline_fp64_pipe": {"bound": "fp64", "achieved": tflops,
                                      "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                                      "frac": tflops / PEAK_FP64_TFLOPS,
                                      "algorithmic_flops_per_launch": flops,
                                      "launch_ms": tk["element_ms"]}}
        if mode == "surrogate":
            res["surrogate_accepted"] = int(c1["surrogate_used"] - c0["surrogate_used"])
            res["surrogate_fallback"] = int(c1["surrogate_fallback"] - c0["surrogate_fallback"])
        results[mode] = res
        dev.close()
        print(f"# {mode}: {ms:.4f} ms/step, {res['value']:.4e} GP/s, {cfg}, transforming "
              f"{res['transforming_fraction']:.3f}, kernels {res['kernel_ms']}",
              file=sys.stderr, flush=True)
    for m in modes:
        results[m]["speedup_vs_fp64"] = results[m]["value"] / results["fp64"]["value"]
    if args.copy_probe_mb > 0:
        # one plain copy of known size: the reference for HBM byte counters
        from alya_sma_amd._lib import check, lib
        from alya_sma_amd.device import DeviceArray
        nb = args.copy_probe_mb << 20
        a, b = DeviceArray(nb // 8, np.float64), DeviceArray(nb // 8, np.float64)
        check(lib().sma_copy_probe(a.ptr, b.ptr, nb, None), "sma_copy_probe")
        synchronize()
        a.free()
        b.free()
    sens = None
    if not args.no_accuracy:
        t0 = time.perf_counter()
        acc, sens = accuracy(prob, modes, start, K, cfgs, sur, args.sur_tol)
        for m in modes:
            results[m].update(acc[m])
        print(f"# accuracy pass {time.perf_counter() - t0:.1f} s: {acc}, fp64 self-sensitivity "
              f"{sens:.3e}", file=sys.stderr, flush=True)
    cpu = None if (args.no_cpu or state0 is None) else cpu_baseline(prob, state0, args.cpu_seconds)
    r64 = results["fp64"]
    out = {
        "metric": METRIC, "value": r64["value"], "unit": "GP-updates/s",
        "n_gpus": 1, "steps": K, "warmup": W, "ms_per_step": r64["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": workload, "mode": "fp64",
                   "window": f"steps {start}..{start + K - 1} of the {cycle}-step load cycle "
                             f"({skip} untimed steps, {W} warm-up), "
                             f"{r64['transforming_fraction']:.3f} of the point updates transforming",
                   "dt": dt, "gauss_points": ngp, "nodes": int(prob["coord"].shape[0])},
        "modes": results,
        # throughput of the FP32 + surrogate step over the FP64 step's
        "surrogate_over_fp64": (results["surrogate"]["speedup_vs_fp64"]
                                if "surrogate" in results else None),
        "fp64_self_sensitivity_trajectory_end": sens,
        "roofline": r64["roofline"], "roofline_fp64_pipe": r64["roofline_fp64_pipe"],
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
