"""Benchmark: Gauss-point updates per second of the explicit SMA step.

Workload (N = 1): the SMA cantilever, hex8 400x20x20 = 160 000 elements
(1.28 M Gauss points) over 50 x 2.5 x 2.5 mm, clamped at x = 0, tip traction
ramped up and released over 5000 central-difference steps, per-element
Young's modulus with seeded +-2% noise.  Inputs are resident in HBM before
the timed region; one step = predictor, element kernel(s), corrector.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--workload cantilever|specimens]

`--workload specimens` runs the batch of 256 uneven SMA specimens (2.7 M
hex8 in one mesh, 2000 steps) instead of the cantilever; `--workload
cylinder` the SMA tube in neo-Hookean tissue (6.05 M hex8, 10000 steps).

Prints one JSON line.  `value` is the throughput of the fastest precision
mode whose stress error stays within 1e-3.  The error of a mode is the
relative L2 difference of the Gauss-point PK2 stress after one step taken
from the FP64 trajectory's state, at ten checkpoints along the run (max
over checkpoints): the error the precision mode adds to a step; the
largest Gauss-point difference relative to the largest stress component is
reported beside it, with the number of points whose branch flags differ from
the FP64 step's (a loading function within rounding of its threshold) and
the largest difference over the points that took the same branch.  The difference between whole trajectories at the last step is reported too,
next to the same difference between two FP64 runs whose initial states
differ by 1e-13 (relative, in u at step 20): transformation fronts make the
undamped load-unload cycle sensitive to any rounding, so the end-of-run
difference measures that sensitivity rather than the precision.

The K timed steps run in ten chunks; two extra eager steps between chunks
(outside the timed sum) time each kernel with HIP events, so `kernel_ms`
and the roofline's `launch_ms` are averages over the whole run.  Every
mode therefore advances W + K + 20 steps.
"""
import argparse
import json
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
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "sld_element_hbm_traffic.txt")
MODES = {"fp64": 0, "fp32": 1, "surrogate": 2}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default: the workload's run)")
    ap.add_argument("--workload", choices=["cantilever", "specimens", "cylinder"], default="cantilever")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--modes", default="fp64,fp32,surrogate")
    ap.add_argument("--sur-tol", type=float, default=1e-4,
                    help="surrogate acceptance: distance in xS to the return map's solution")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the error pass")
    ap.add_argument("--tune", choices=["auto", "none"], default="auto",
                    help="auto: run the library autotuner after the warm-up steps")
    ap.add_argument("--assembly", choices=["gather", "scatter"], default=None,
                    help="fix the assembly (disables the autotuner)")
    ap.add_argument("--epb", type=int, default=0, help="elements per block (disables the autotuner)")
    ap.add_argument("--copy-probe-mb", type=int, default=256,
                    help="size of the calibration copy kernel run (HBM counter reference)")
    return ap.parse_args()


def kernel_name(mode, cfg):
    r = "double" if mode == "fp64" else "float"
    sur = 2 if mode == "surrogate" else 0
    ga = "true" if cfg["assembly"] == "gather" else "false"
    # the last template argument: 0 = each element's own law (2 = the
    # neo-Hookean-only launch that follows it on a law-grouped mesh)
    return f"sld_element<{r}, {cfg['elems_per_block']}, false, {sur}, {ga}, 0>"


def algorithmic_bytes(prob, dev, mode, assembly):
    """Bytes one element launch must move at least (average over launches).

    Per Gauss point: connectivity 4 B, the element's bulk and shear moduli
    2 B, and on SMA elements the history read + written 48 B (neo-Hookean
    elements have none).  Gather assembly: per Gauss point the corner's slot
    index 4 B and its force written (3 x 8 B in FP64, 3 x 4 B in FP32), per
    node coordinates + displacement read 48 B.  Coloured scatter: per node
    the launch touches coordinates + displacement 48 B and f_int
    read-modify-write (2 x 3 x 8 B or 2 x 3 x 4 B)."""
    from alya_sma_amd.problems import LAW_SMA
    fb = 8 if mode == "fp64" else 4
    lnods = prob["lnods"]
    ne = lnods.shape[0]
    per_gp = np.full(ne, 6.0 + 48.0)
    if prob.get("law_of_mat") is not None:
        mat = prob.get("mat_of_elem")
        law = np.asarray(prob["law_of_mat"])[mat if mat is not None else np.zeros(ne, np.int64)]
        per_gp = np.where(law == LAW_SMA, 6.0 + 48.0, 6.0)
    if assembly == "gather":
        npn = prob["coord"].shape[0]
        return float(8 * per_gp.sum() + ne * 8 * (4 + 3 * fb) + npn * 48)
    ptr, order = dev.coloring()
    per = []
    for c in range(len(ptr) - 1):
        els = order[ptr[c]:ptr[c + 1]]
        npc = len(np.unique(lnods[els].ravel()))
        per.append(8 * per_gp[els].sum() + npc * (48 + 6 * fb))
    return float(np.mean(per))


def measured_traffic(workload, kernel):
    """HBM bytes per element phase of `kernel` on `workload` from the committed
    rocprofv3 PMC summary (FETCH_SIZE and WRITE_SIZE scaled by a plain copy
    kernel): on a law-grouped mesh the phase is two launches, `kernel` over
    the SMA range (last template argument 0) and its neo-Hookean-only
    variant (2), whose bytes add.  None when the pair was not measured."""
    nh = kernel[:-2] + "2>"
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


def cpu_baseline(prob, budget_s):
    """Time the FP64 CPU oracle (OpenMP over the elements of each colour) on
    the first steps of the same workload."""
    from oracle import oracle_py as O
    from alya_sma_amd import _lib
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    lnods = prob["lnods"]
    color = np.zeros(lnods.shape[0], dtype=np.int32)
    nc = _lib.lib().sld_color_elements(lnods.shape[0], prob["coord"].shape[0],
                                       lnods.ctypes.data, color.ctypes.data)
    order = np.argsort(color, kind="stable").astype(np.int32)
    ptr = np.concatenate([[0], np.cumsum(np.bincount(color, minlength=nc))]).astype(np.int32)
    orc = O.StepOracle(prob, elem_order=order, color_ptr=ptr, keep_stress=False)
    dt = prob["run"]["dt"]
    orc.step(dt, scheme=0)                     # first step outside the sample (page-in)
    n, t0 = 0, time.perf_counter()
    while True:
        orc.step(dt, scheme=0)
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s or n >= 200:
            break
    ngp = prob["lnods"].shape[0] * 8
    cpu = "unknown"
    try:
        cpu = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        cpu = [l.split(":", 1)[1].strip() for l in cpu.splitlines() if l.startswith("Model name")][0]
    except Exception:
        pass
    return {"value": ngp * n / el, "unit": "GP-updates/s", "cores": threads, "kind": "port",
            "sample": f"steps 2..{n + 1} of the same workload, FP64 C oracle "
                      f"(oracle/sld_step.c, gcc -O3, OpenMP over the elements of each colour), "
                      f"{el:.1f} s wall on {threads} threads of {cpu}"}


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


def accuracy(prob, modes, K, cfgs, sur=None, tol=1e-4):
    """One-step stress error at checkpoints, trajectory error at the end,
    FP64 self-sensitivity (see the module docstring)."""
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
    first = min(20, K)
    for d in [ref, pert] + list(traj.values()):
        d.step(dt, nsteps=first)
    pert.set("u", pert.get("u") * (1.0 + 1e-13))
    done, one, onemax = first, {m: 0.0 for m in others}, {m: 0.0 for m in others}
    # points whose branch flags differ from the FP64 step's, and the largest
    # difference over the points that took the same branch
    flips, same_max = {m: 0 for m in others}, {m: 0.0 for m in others}
    chunk = max(2, (K - first) // 10)
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
        print(f"# accuracy pass: {done} of {K} steps", file=sys.stderr, flush=True)
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
    from alya_sma_amd.solid import Solidz

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
    nelem = prob["lnods"].shape[0]
    ngp = nelem * 8
    K = args.steps if args.steps is not None else prob["run"]["nsteps"]
    W = args.warmup
    modes = [m for m in args.modes.split(",") if m]
    tune = args.tune == "auto" and not args.assembly and not args.epb
    results, cfgs, roofs = {}, {}, {}
    sur = surrogate_network() if "surrogate" in modes else None
    for mode in modes:
        dev = make_device(prob, mode, sur, args.sur_tol)
        if args.assembly:
            dev.set_assembly(args.assembly)
        if args.epb:
            dev.set_tile(args.epb)
        dev.step(dt, nsteps=W)
        cfg = {"assembly": args.assembly or "gather", "elems_per_block": args.epb or 8,
               "xcd_remap": 1, "persistent": 0}
        if tune:
            cfg = dev.autotune(dt, nsteps=3)
            cfg.pop("ms_per_step")
        cfgs[mode] = cfg
        # K timed steps in ten chunks, each bracketed by synchronisations;
        # between chunks (outside the timed sum) one eager step with HIP
        # events samples the per-launch kernel times, so their average
        # covers the elastic and the transforming phases of the run
        el, samples = 0.0, []
        nchunk = min(10, K)
        for c in range(nchunk):
            n = K // nchunk + (1 if c < K % nchunk else 0)
            dev.sync()
            synchronize()
            t0 = time.perf_counter()
            dev.step(dt, nsteps=n)
            dev.sync()
            synchronize()
            el += time.perf_counter() - t0
            samples.append(dev.time_kernels(dt, nsteps=2))
        ms = 1e3 * el / K
        tk = {k: float(np.mean([x[k] for x in samples])) for k in samples[0]}
        cnt = dev.counters()
        kname = kernel_name(mode, cfg)
        nbytes = algorithmic_bytes(prob, dev, mode, cfg["assembly"])
        achieved = nbytes / (tk["element_ms"] * 1e-3) / 1e9
        roofs[mode] = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                       "frac": achieved / PEAK_HBM_GBS,
                       "traffic": measured_traffic(args.workload, kname),
                       "kernel": kname, "algorithmic_bytes_per_launch": nbytes,
                       "launch_ms": tk["element_ms"]}
        results[mode] = {"value": ngp / (ms * 1e-3), "ms_per_step": ms, "launch": cfg,
                         "kernel_ms": {k: round(v, 5) for k, v in tk.items()},
                         "transforming_gp_updates": cnt["kf_AS"] + cnt["kf_SA"]}
        if mode == "surrogate":
            results[mode]["surrogate_accepted"] = cnt["surrogate_used"]
            results[mode]["surrogate_fallback"] = cnt["surrogate_fallback"]
        dev.close()
        print(f"# {mode}: {ms:.4f} ms/step, {results[mode]['value']:.4e} GP/s, {cfg}, kernels "
              f"{results[mode]['kernel_ms']}", file=sys.stderr, flush=True)
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
        acc, sens = accuracy(prob, modes, K + W + 2 * min(10, K), cfgs, sur, args.sur_tol)
        for m in modes:
            results[m].update(acc[m])
        print(f"# accuracy pass {time.perf_counter() - t0:.1f} s: {acc}, fp64 self-sensitivity "
              f"{sens:.3e}", file=sys.stderr, flush=True)
    ok = [m for m in modes if results[m].get("stress_rel_l2_one_step", 0.0) <= 1e-3]
    best = max(ok, key=lambda m: results[m]["value"]) if ok else "fp64"
    cpu = None if args.no_cpu else cpu_baseline(prob, args.cpu_seconds)
    out = {
        "metric": METRIC, "value": results[best]["value"], "unit": "GP-updates/s",
        "n_gpus": 1, "steps": K, "warmup": W, "ms_per_step": results[best]["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32+f64", "surrogate": "bf16+f32+f64"}[best],
        "data": "synthetic",
        "config": {"workload": workload, "mode": best, "dt": dt,
                   "gauss_points": ngp, "nodes": int(prob["coord"].shape[0])},
        "modes": results,
        "fp64_self_sensitivity_trajectory_end": sens,
        "roofline": roofs[best], "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
