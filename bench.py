#!/usr/bin/env python3
"""Benchmark: Irving-Kirkwood stress fields, frames per second on one MI355X.

Workload (one "step"): 100 frames of a 2D polydisperse12 glass, N = 65,536
particles each (seeded lattice+jitter recipe, pyglasstools_amd/workload.py,
seeds 2000..2099), all resident in HBM; FP64 virial + kinetic stress on a
256 x 256 grid, Octic kernel r_cg = 2.0, 5-node Gauss-Legendre bond
integral.  value = frames / second over K timed steps.

Printed as one JSON line.  `roofline` is the field-gather kernel (the
dominant kernel): algorithmic FP64 flops per launch (112 per pair-grid
contribution U, 23 per kinetic (grid point, particle) pair, counted by the
kernel itself) divided by its average launch time from HIP events on the
engine's stream.  `cpu_baseline` is the FP64 CPU oracle (a C restatement of
the reference loop) on the host cores over a bounded sample of the same
frames.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "IK stress frames/sec at ≤1e-5 rel err vs FP64; surrogate vs plain quad"
FLOP_U_2D, FLOP_KIN_2D = 112, 23
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def make_frames(nframes, n, config=2):
    from pyglasstools_amd import workload

    systems = [workload.make_system(n, 2, 1000 * config + f) for f in range(nframes)]
    return systems


def traffic_from_profile(mode, workload="c2"):
    """HBM bytes per launch of the field gather of the given kind ("tile" or
    "lane") on the given workload ("c2": 2D N = 65,536; "c3": 3D N = 131,072
    FP32 form) measured by rocprofv3 PMC counters (FETCH_SIZE and
    WRITE_SIZE, calibrated on a device copy), with the frames per launch they
    were measured at, if a committed summary exists
    (profiles/gather_<workload>_<mode>_traffic.json)."""
    p = os.path.join(ROOT, "profiles", "gather_%s_%s_traffic.json" % (workload, mode))
    if os.path.exists(p):
        try:
            with open(p) as f:
                d = json.load(f)
            return d.get("bytes_per_launch"), d.get("frames_per_launch")
        except (OSError, ValueError):
            pass
    return None, None


def host_cpu():
    """CPU model, logical CPUs online, and the CPUs this process may use (its
    affinity mask: the host's share given to this job)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"model": model, "cpus_online": os.cpu_count(), "cpus_usable": usable}


def cpu_baseline(systems, grid, max_rcut, seconds=10.0):
    """Oracle on the host: after a warm-up frame, frames in order until
    `seconds` of wall time, the rate from the median per-frame time, on
    every CPU this process may use (the affinity mask; an OMP_NUM_THREADS set
    by the host environment caps it), built for this host's CPU
    (-march=native, as the reference's CompilerFlagsSetup.cmake:14 builds)."""
    import tempfile

    import oracle

    info = host_cpu()
    threads = info["cpus_usable"]
    env_threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    if env_threads > 0:
        threads = min(threads, env_threads)
    build = "-O3 -march=native -fopenmp (built on this host)"
    try:
        oracle.use_library(oracle.build_native(tempfile.mkdtemp(prefix="pgt_oracle_")))
    except Exception as exc:  # no compiler on this host: the portable build
        build = "prebuilt -O3 -march=x86-64-v3 -fopenmp (native build failed: %s)" % exc
    out = {}
    for mode, name in ((1, "clean"), (0, "faithful")):
        # one warm-up frame, then frames in order until `seconds`; the rate is
        # one over the median per-frame time (the mean beside it)
        oracle.ik_compute(systems[0], grid, fields=("virial", "kinetic"), max_rcut=max_rcut, mode=mode,
                          nthreads=threads)
        done, ts, t0 = 0, [], time.perf_counter()
        while done < len(systems):
            t1 = time.perf_counter()
            oracle.ik_compute(systems[done], grid, fields=("virial", "kinetic"), max_rcut=max_rcut, mode=mode,
                              nthreads=threads)
            ts.append(time.perf_counter() - t1)
            done += 1
            if time.perf_counter() - t0 >= seconds:
                break
        dt = time.perf_counter() - t0
        out[name] = {"frames": done, "seconds": dt, "fps": 1.0 / float(np.median(ts)), "fps_mean": done / dt}
    info["build"] = build
    info["threads"] = threads
    return out, threads, info


DRIVER_SCRIPT = """
import json, sys, time
t_start = time.perf_counter()
sys.path.insert(0, {root!r})
import pyglasstools as pglass
import pyglasstools.io as io
import pyglasstools.utils as utils
import pyglasstools.potential as pot
import pyglasstools.irvingkirkwood as ik
import pyglasstools.irvingkirkwood.cgfunc as cgfunc
poly12 = pot.polydisperse12(v0=1.0, eps=0.2, rcut=1.25)
mycgfunc = cgfunc.octic(rcut=2.0, mode="fixedpoint", order=5)
mysystem = utils.read_gsd(filename={gsd!r}, checkpointfile="restart.log")
MyCalculator = ik.ikcalculator(cgfunc=mycgfunc, dx=0.25)
names = ['virialstress_00', 'virialstress_11', 'virialstress_01']
mylogger = io.fieldlogger(keyword="test-dump", names=names, solver=MyCalculator)
t_ready = time.perf_counter()
acc = {{"update": 0.0, "run": 0.0, "save": 0.0}}
runs = []
def timed(obj, name, key):
    f = getattr(obj, name)
    def g(*a, **k):
        t = time.perf_counter()
        r = f(*a, **k)
        dt = time.perf_counter() - t
        acc[key] += dt
        if key == "run":
            runs.append(dt)
        return r
    setattr(obj, name, g)
timed(MyCalculator, "update", "update")
timed(MyCalculator, "run", "run")
timed(mylogger, "save", "save")
t0 = time.perf_counter()
pglass.analyze(frame_list=range(180, 200))
t1 = time.perf_counter()
print("DRIVER " + json.dumps({{"frames": 20, "analyze_s": t1 - t0, "setup_s": t_ready - t_start,
                              "script_s": t1 - t_start, "update_s": acc["update"], "compute_s": acc["run"],
                              "save_s": acc["save"], "run_s_each": runs}}))
"""


def driver_script_time():
    """The reference's IK example sequence (example_scripts/irvingkirkwood/
    test-irvingkirkwood-module.py: potential, kernel, GSD ingest,
    ikcalculator, fieldlogger, analyze) on the committed fixture, frames
    180..199 (1024 particles, 127 x 127 grid), run as a script through the
    pybind11 surface: wall time of the whole script and of its parts."""
    import subprocess
    import tempfile

    gsd = os.path.join(ROOT, "tests", "golden", "polydisperse12.gsd")
    with tempfile.TemporaryDirectory(prefix="pgt_driver_") as d:
        with open(os.path.join(d, "drive.py"), "w") as f:
            f.write(DRIVER_SCRIPT.format(root=ROOT, gsd=gsd))
        t0 = time.perf_counter()
        r = subprocess.run([sys.executable, "drive.py"], cwd=d, capture_output=True, text=True, timeout=600)
        wall = time.perf_counter() - t0
    if r.returncode != 0:
        return {"error": r.stderr[-500:]}
    line = [x for x in r.stdout.splitlines() if x.startswith("DRIVER ")][-1]
    out = json.loads(line[len("DRIVER "):])
    out["process_wall_s"] = wall
    runs = out.pop("run_s_each")
    out["compute_ms_per_frame"] = out["compute_s"] / out["frames"] * 1e3
    # the first compute creates the engine (HIP runtime, allocations)
    out["compute_first_frame_ms"] = runs[0] * 1e3
    out["compute_ms_per_frame_after_first"] = float(np.median(runs[1:])) * 1e3
    return out


def parity_sample(systems, grid, max_rcut, got_virial, got_kinetic, stride=16):
    """Oracle vs the GPU result of frame 0 on every `stride`-th grid point."""
    import oracle

    sub = np.arange(0, len(grid), stride)
    ref = oracle.ik_compute(systems[0], grid[sub], fields=("virial", "kinetic"), max_rcut=max_rcut, with_abs=True)
    scale = np.maximum(np.abs(ref["virial"]), 1e-6 * ref["virial_abs"])
    ev = float((np.abs(got_virial[sub] - ref["virial"]) / np.where(scale > 0, scale, 1)).max())
    ek = float(np.abs(got_kinetic[sub] - ref["kinetic"]).max() / max(np.abs(ref["kinetic"]).max(), 1e-300))
    rel = float(np.abs(got_virial[sub] - ref["virial"]).max() / np.abs(ref["virial"]).max())
    return {"grid_points": int(len(sub)), "virial_max_scaled_err": ev, "virial_max_err_over_max": rel,
            "kinetic_max_err_over_max": ek}


FLOP_U_3D = 142
PEAK_FP32_TFLOPS = 157.3


def precision_form(args):
    """The FP32 precision form on the 3D workload: polydisperse12, N = 131,072
    at rho = 1 (seeds 3000..3019), 20 frames resident in HBM, virial stress on
    a 64^3 grid, Octic r_cg = 2.0, 5-node rule.  Pair forces and kernel
    weights in FP32; minimum-image displacements and the per-grid-point sums
    in FP64.  Reports frames/s, contributions/s, the gather's share of the
    FP32 VALU peak, and the per-component error against the FP64 oracle on a
    sample of grid points of frame 0 (with the FP64 GPU form beside it)."""
    import torch

    import oracle
    from pyglasstools_amd import engine, workload

    dev = torch.device("cuda:0")
    F, N, NG = 20, 131072, 64
    log("generating %d 3D frames of N=%d" % (F, N))
    systems = [workload.make_system(N, 3, 3000 + f) for f in range(F)]
    L = systems[0]["L"]
    grid = workload.cartesian_grid_3d(L[0], NG)
    G = len(grid)
    d0 = float(systems[0]["diam"].max())
    max_rcut = max(1.25 * d0, d0)
    pos = torch.tensor(np.stack([s["pos"] for s in systems]), dtype=torch.float64, device=dev)
    diam = torch.tensor(np.stack([s["diam"] for s in systems]), dtype=torch.float64, device=dev)
    out_v = torch.empty((F, G, 3, 3), dtype=torch.float64, device=dev)
    eng = engine.Engine(dim=3, max_rcut=max_rcut, precision="fp32")
    eng.set_box(L, [1, 1, 1])
    eng.set_grid(grid)
    eng.set_tuning(frames_per_launch=F)

    def step():
        eng.compute_frames(F, N, pos, diam, None, None, None, fields=("virial",), d_virial=out_v)

    from pyglasstools_amd.autotune import Autotuner

    if args.block and args.pl:
        mode = "tile" if args.gather == "auto" else args.gather
        eng.set_gather_mode(mode)
        eng.set_tuning(args.block, F, args.pl, args.reach)
        tuned = {"gather": mode, "threads_per_block": args.block, "frames_per_launch": F,
                 "points_per_lane": args.pl, "cell_reach": args.reach}
    else:
        log("autotuning gather kind x threads per workgroup x points per lane")
        gathers = ("lane", "tile") if args.gather == "auto" else (args.gather,)
        tuned, _ = Autotuner().tune(eng, step, ("c3", N, G, F, args.gather), F, fpl=(F,),
                                    sync=torch.cuda.synchronize, gathers=gathers)
        log("tuned: %s" % tuned)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    gather_ms, n_contrib = 0.0, 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        st = eng.stats()
        gather_ms += st["gather_ms"]
        n_contrib += st["n_contrib"]
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    fps = F * args.steps / elapsed
    log("precision form: %d steps in %.3f s: %.2f frames/s" % (args.steps, elapsed, fps))
    avg_launch_s = gather_ms / 1e3 / args.steps
    flops = FLOP_U_3D * n_contrib / args.steps
    achieved = flops / avg_launch_s / 1e12
    traffic, traffic_fpl = traffic_from_profile(tuned["gather"], "c3")
    if traffic is not None and traffic_fpl != F:
        traffic = traffic * F / traffic_fpl  # per-launch bytes scale with frames per launch

    # error vs the FP64 oracle on a sample of grid points of frame 0; the FP64
    # GPU form on the same points for comparison
    v32 = out_v[0].cpu().numpy()
    e64 = engine.Engine(dim=3, max_rcut=max_rcut, precision="fp64")
    e64.set_system(systems[0], grid)
    v64 = e64.compute(fields=("virial",))["virial"]
    rng = np.random.default_rng(7)
    idx = np.sort(rng.choice(G, size=min(G, args.c3_sample), replace=False))
    threads = min(int(os.environ.get("OMP_NUM_THREADS", "0") or os.cpu_count() or 1), 16)
    ref = oracle.ik_compute(systems[0], grid[idx], max_rcut=max_rcut, fields=("virial",), nthreads=threads)["virial"]
    comps = {}
    for a in range(3):
        for b in range(a, 3):
            scale = np.abs(ref[:, b, a]).max()
            comps["%d%d" % (a, b)] = {
                "fp32_form_max_abs_err_over_max_abs": float(np.abs(v32[idx, b, a] - ref[:, b, a]).max() / scale),
                "fp64_form_max_abs_err_over_max_abs": float(np.abs(v64[idx, b, a] - ref[:, b, a]).max() / scale)}
    worst = max(c["fp32_form_max_abs_err_over_max_abs"] for c in comps.values())
    log("precision form: worst component error %.3e" % worst)
    line = {"metric": METRIC, "value": round(fps, 3), "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 4), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
            "data": "synthetic (seeded lattice+jitter polydisperse glass, 20 frames resident in HBM)",
            "config": {"workload": "3D polydisperse12 N=131072, 20 frames, 64^3 grid, Octic r_cg=2.0, GL5, FP32 "
                                   "forces and kernel weights, FP64 minimum image and accumulation, virial stress",
                       "frames_per_step": F, "n_particles": N, "grid": "64^3", "parallelism": "single GPU"},
            "roofline": {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
                         "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": traffic,
                         "kernel": ("k_gather_tile" if tuned["gather"] == "tile" else "k_gather") +
                                   " (field gather, FP32 form)",
                         "compute": "FP32 vector ALU (published MI355X FP32 peak 157.3 TFLOP/s, which counts packed "
                                    "two-wide FP32; this build issues unpacked FP32, whose rate equals the FP64 "
                                    "rate of 78.6 TFLOP/s)",
                         "frac_of_unpacked_rate": round(achieved / PEAK_FP64_TFLOPS, 4),
                         "avg_launch_ms": round(avg_launch_s * 1e3, 4), "flops_per_launch": flops},
            "cpu_baseline": None,
            "extra": {"contributions_per_s": n_contrib / elapsed, "contributions_per_frame": n_contrib / (F * args.steps),
                      "tuning": tuned,
                      "error_vs_fp64_oracle": {"grid_points_sampled": int(len(idx)), "frame": 0,
                                               "per_component": comps, "worst_fp32": worst}}}
    if traffic is not None:
        # HBM bytes of the gather (PMC counters) over its launch time, against
        # 8 TB/s; algorithmic: positions and diameters in, the field out
        rl = line["roofline"]
        rl["hbm_gbs"] = round(traffic / avg_launch_s / 1e9, 1)
        rl["hbm_frac"] = round(traffic / avg_launch_s / 1e9 / PEAK_HBM_GBS, 4)
        rl["algorithmic_bytes_per_launch"] = F * (N * (24 + 8) + G * 72)
    print(json.dumps(line), flush=True)


PEAK_BF16_TFLOPS = 2500.0  # dense BF16 matrix-core peak
FLOP_MLP = 2 * (5 * 64 + 2 * 64 * 64 + 64)  # 17,152 per network evaluation


def surrogate_comparison(epochs, steps, log_prefix="surrogate"):
    """The learned bond-integral surrogate against plain quadrature on the
    config-1 workload (2D polydisperse12 N=2000, 32 x 32 grid, Octic r_cg =
    2.0, 5-node rule), with three disjoint sets of config-1-recipe frames:
    the network trains (FP32 Adam on the engine's own kernels) on every
    (grid point, bond) sample of frames 0..19, its per-class error bound is
    calibrated on frame 20, and the virial field is evaluated on frame 21 in
    the plain FP64 form and in the surrogate form at the metric's tolerance
    1e-5 and at the calibrated inside-class bound: frames/s (host API, one
    frame per call), the share of pairs the network serves, and the largest
    component error against the FP64 oracle.  The bound is empirical (twice
    the largest calibration error) and per bond integral, not a field
    tolerance.  Returns (result dict, trained params, training samples)."""
    import torch

    import oracle
    from pyglasstools_amd import engine, surrogate, workload

    dev = torch.device("cuda:0")
    frames = [workload.config_frame(1, f)[0] for f in range(22)]
    grid = workload.config_frame(1, 0)[1]
    d0 = float(frames[0]["diam"].max())
    max_rcut = max(1.25 * d0, d0)  # Calculator.h:41-45, from the first frame
    e = engine.Engine(dim=2, max_rcut=max_rcut)
    t0 = time.perf_counter()
    x, t = surrogate.collect_samples(e, frames[:20], grid, dev)
    xc, tc = surrogate.collect_samples(e, frames[20:21], grid, dev)
    t_samples = time.perf_counter() - t0
    log("%s: %d training samples, %d calibration samples (%.2f s)" % (log_prefix, x.shape[0], xc.shape[0], t_samples))
    t0 = time.perf_counter()
    p, hist = surrogate.train(x, t, epochs=epochs, batch=1024, lr=1e-3, lr_final=1e-5, seed=0, log=log)
    t_train = time.perf_counter() - t0
    untrained = torch.tensor(surrogate.init_params(0), device=dev)
    err, err_un = surrogate.errors(p, xc, tc), surrogate.errors(untrained, xc, tc)
    ins = surrogate.inside_class(xc)
    bounds = surrogate.calibrate(p, xc, tc)
    log("%s: trained %d epochs in %.1f s, calibration mean |err| %.3e (untrained %.3e), bounds %s" %
        (log_prefix, epochs, t_train, err.mean().item(), err_un.mean().item(), bounds))

    fr = frames[21]
    e.set_system(fr, grid)
    ref = oracle.ik_compute(fr, grid, max_rcut=max_rcut, fields=("virial",), with_abs=True)["virial"]

    def rate(fn):
        # per-call times; the median keeps one host stall out of the rate
        fn()
        ts = []
        for _ in range(max(steps, 50)):
            t1 = time.perf_counter()
            out = fn()
            ts.append(time.perf_counter() - t1)
        return 1.0 / float(np.median(ts)), out, len(ts) / float(np.sum(ts))

    def comp_err(v):
        return max(float(np.abs(v[:, b, a] - ref[:, b, a]).max() / np.abs(ref[:, b, a]).max())
                   for a, b in ((0, 0), (1, 1), (0, 1)))

    fps_plain, v_plain, fps_plain_mean = rate(lambda: e.compute(fields=("virial",))["virial"])
    forms = {}
    for name, tol in (("tol_1e-5", 1e-5), ("tol_inside_bound", bounds["inside"])):
        fps, v, fps_mean = rate(lambda: e.compute_surrogate(p, bounds["inside"], bounds["crossing"], tol))
        st = e.stats()
        forms[name] = {"tolerance_per_integral": tol, "frames_per_s": fps, "frames_per_s_mean": fps_mean,
                       "speed_vs_plain": fps / fps_plain, "network_share": st["n_network"] / st["n_pairs"],
                       "pairs": st["n_pairs"], "max_component_err_over_max_vs_fp64_oracle": comp_err(v),
                       "pair_pipeline_ms": st["surrogate_ms"], "device_ms_per_call": st["total_ms"]}
        log("%s form %s: %.1f frames/s (plain %.1f), network share %.3f, error %.3e" %
            (log_prefix, name, fps, fps_plain, forms[name]["network_share"],
             forms[name]["max_component_err_over_max_vs_fp64_oracle"]))
    res = {"workload": "2D polydisperse12 N=2000, 32x32 grid, Octic r_cg=2.0, GL5, virial; train frames 0-19, "
                       "calibrate frame 20, evaluate frame 21 (config-1 recipe)",
           "plain_fp64_frames_per_s": fps_plain, "plain_fp64_frames_per_s_mean": fps_plain_mean,
           "plain_fp64_max_component_err_over_max": comp_err(v_plain),
           "timing": "host API, one frame per call, median of the per-call times (mean beside it)",
           "surrogate_forms": forms, "epochs": epochs, "samples": int(x.shape[0]),
           "calibration_samples": int(xc.shape[0]), "sample_seconds": t_samples, "train_seconds": t_train,
           "train_history": hist, "calibration_mean_abs_err": err.mean().item(),
           "calibration_max_abs_err": err.max().item(),
           "calibration_inside_max_abs_err": err[ins].max().item() if bool(ins.any()) else None,
           "untrained_calibration_mean_abs_err": err_un.mean().item(),
           "error_bounds_empirical_per_integral": bounds}
    return res, p, x


def learned_surrogate(args):
    """Config 4: the surrogate comparison (surrogate_comparison) with the full
    training schedule, and the BF16 matrix-core inference rate over the
    training samples (roofline against the BF16 peak)."""
    import torch

    from pyglasstools_amd import engine

    res, p, x = surrogate_comparison(args.c4_epochs, args.steps)
    y = torch.empty((x.shape[0],), dtype=torch.float32, device=x.device)
    engine.mlp_infer(p, x, y)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 10
    ev0.record()
    for _ in range(reps):
        engine.mlp_infer(p, x, y)
    ev1.record()
    torch.cuda.synchronize()
    infer_s = ev0.elapsed_time(ev1) / 1e3 / reps
    infer_tflops = FLOP_MLP * x.shape[0] / infer_s / 1e12
    main = res["surrogate_forms"]["tol_1e-5"]
    res.update({"inference_samples_per_s": x.shape[0] / infer_s, "quadrature_flops_per_integral": 82})
    line = {"metric": METRIC, "value": round(main["frames_per_s"], 3), "unit": "frames/s", "n_gpus": 1,
            "steps": args.steps, "warmup": 1, "ms_per_step": round(1e3 / main["frames_per_s"], 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16+f64",
            "data": "synthetic (seeded config-1 recipe frames; network trained here on their quadrature samples)",
            "config": {"workload": "2D polydisperse12 N=2000, 1 frame, 32x32 grid, Octic r_cg=2.0; bond integral "
                                   "from a 5-64-64-64-1 tanh MLP (BF16 matrix cores, FP32 accumulate) where its "
                                   "calibrated error bound is within 1e-5, 5-node quadrature elsewhere",
                       "parallelism": "single GPU"},
            "roofline": {"bound": "mfma", "achieved": round(infer_tflops, 3), "peak": PEAK_BF16_TFLOPS,
                         "unit": "TFLOP/s", "frac": round(infer_tflops / PEAK_BF16_TFLOPS, 5), "traffic": None,
                         "kernel": "k_mlp_infer (BF16 MFMA network inference over %d samples)" % x.shape[0],
                         "flops_per_sample": FLOP_MLP, "seconds_per_launch": infer_s},
            "cpu_baseline": None,
            "extra": res}
    print(json.dumps(line), flush=True)


C5_N = (4096, 8192, 16384, 24576, 32768, 49152, 65536, 98304)


def gs_nodes(graph):
    """Nodes of a captured torch graph made with keep_graph=True, before it
    is instantiated (hipGraphGetNodes; None where the runtime does not tell)."""
    try:
        import ctypes

        hip = ctypes.CDLL("libamdhip64.so")
        g = ctypes.c_void_p(graph.raw_cuda_graph())
        n = ctypes.c_size_t(0)
        if not g.value or hip.hipGraphGetNodes(g, None, ctypes.byref(n)) != 0:
            return None
        return int(n.value)
    except Exception:
        return None


def sweep_roofline(r):
    """Roofline of the sweep's FP64 field gather (one launch over every
    system's frames): canonical flops of its contributions and ball pairs over
    its HIP-event time in an eager run."""
    t = r["sweep_gather_ms_eager"] / 1e3
    achieved = r["sweep_gather_flops"] / t / 1e12
    return {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": None,
            "kernel": "k_sweep_gather (tile field gather over all systems' frames)",
            "compute": "FP64 vector ALU (published MI355X FP64 peak 78.6 TFLOP/s)",
            "avg_launch_ms": round(t * 1e3, 4), "flops_per_launch": r["sweep_gather_flops"]}


def batched_sweep(args):
    """Serving-style sweep: 1000 2D frames from 8 systems of uneven N (125
    frames each, seeds 5000 + 1000 k + f), each on a 128 x 128 grid over its
    own box, virial + kinetic.  One engine per system, launch shape tuned per
    system by the autotuner; then the whole sweep runs (a) eagerly, system
    after system, (b) as one HIP graph holding every system's launches
    (pgt_ik_enqueue_frames captured on one stream), replayed, (c) the same
    with the systems on streams forked inside the graph, and (d) as one
    sweep (pgt_ik_sweep): one launch per stage over every system's frames,
    eagerly and captured.  Both for the FP64 plain form and the FP32 form;
    the line's value is (d) captured, FP64."""
    import torch

    from pyglasstools_amd import engine, workload
    from pyglasstools_amd.autotune import Autotuner

    dev = torch.device("cuda:0")
    FR, NG = 125, 128
    log("generating %d systems x %d frames" % (len(C5_N), FR))
    systems = []
    for k, N in enumerate(C5_N):
        frames = [workload.make_system(N, 2, 5000 + 1000 * k + f) for f in range(FR)]
        L = frames[0]["L"]
        d0 = float(frames[0]["diam"].max())
        t = lambda key: torch.tensor(np.stack([fr[key] for fr in frames]), dtype=torch.float64, device=dev)
        systems.append({"N": N, "L": L, "max_rcut": max(1.25 * d0, d0), "grid": workload.cartesian_grid_2d(L[0], NG),
                        "pos": t("pos"), "diam": t("diam"), "mass": t("mass"), "vel": t("vel")})
        del frames
    G = NG * NG
    results = {}
    for prec in ("fp64", "fp32"):
        tuner = Autotuner()
        jobs = []
        for sy in systems:
            e = engine.Engine(dim=2, max_rcut=sy["max_rcut"], precision=prec)
            e.set_box(sy["L"], [1, 1, 0])
            e.set_grid(sy["grid"])
            ov = torch.empty((FR, G, 2, 2), dtype=torch.float64, device=dev)
            ok = torch.empty_like(ov)

            def run(e=e, sy=sy, ov=ov, ok=ok):
                e.compute_frames(FR, sy["N"], sy["pos"], sy["diam"], sy["mass"], sy["vel"], None,
                                 fields=("virial", "kinetic"), d_virial=ov, d_kinetic=ok)

            best, _ = tuner.tune(e, run, ("c5", prec, sy["N"]), FR, blocks=(128, 256), fpl=(25, FR),
                                 sync=torch.cuda.synchronize, repeats=1)
            jobs.append({"e": e, "sy": sy, "ov": ov, "ok": ok, "run": run, "tuning": best})
        # (a) eager
        for j in jobs:
            j["run"]()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        nc = 0
        for _ in range(args.steps):
            for j in jobs:
                j["run"]()
                nc += j["e"].stats()["n_contrib"]
        torch.cuda.synchronize()
        t_eager = (time.perf_counter() - t0) / args.steps
        # (b) one graph over all systems
        s = torch.cuda.Stream()
        for j in jobs:
            j["e"].set_stream(s.cuda_stream)
        with torch.cuda.stream(s):
            for j in jobs:
                j["run"]()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            for j in jobs:
                sy = j["sy"]
                j["e"].enqueue_frames(FR, sy["N"], sy["pos"], sy["diam"], sy["mass"], sy["vel"], None,
                                      fields=("virial", "kinetic"), d_virial=j["ov"], d_kinetic=j["ok"])
        g.replay()
        torch.cuda.synchronize()
        for j in jobs:
            j["e"].finish()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            g.replay()
        torch.cuda.synchronize()
        t_graph = (time.perf_counter() - t0) / args.steps
        nc_graph = 0
        for j in jobs:
            j["e"].finish()  # also checks the bond capacity over every replay
            nc_graph += j["e"].stats()["n_contrib"]
        # (c) one graph whose systems run on K streams forked from the capture
        # stream and joined back (the largest system paired with the smallest)
        K = args.sweep_streams
        side = [torch.cuda.Stream() for _ in range(K)]
        order = sorted(range(len(jobs)), key=lambda i: jobs[i]["sy"]["N"])
        lane = {}
        for k in range(len(order)):
            lane[order[k]] = k if k < K else 2 * K - 1 - k
        for i, j in enumerate(jobs):
            j["e"].set_stream(side[lane[i] % K].cuda_stream)
            with torch.cuda.stream(side[lane[i] % K]):
                j["run"]()
        torch.cuda.synchronize()
        g2 = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g2, stream=s):
            for st_ in side:
                st_.wait_stream(s)
            for j in jobs:
                sy = j["sy"]
                j["e"].enqueue_frames(FR, sy["N"], sy["pos"], sy["diam"], sy["mass"], sy["vel"], None,
                                      fields=("virial", "kinetic"), d_virial=j["ov"], d_kinetic=j["ok"])
            for st_ in side:
                s.wait_stream(st_)
        g2.replay()
        torch.cuda.synchronize()
        for j in jobs:
            j["e"].finish()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            g2.replay()
        torch.cuda.synchronize()
        t_graph_k = (time.perf_counter() - t0) / args.steps
        for j in jobs:
            j["e"].finish()
            j["e"].set_stream(0)
        # (d) one sweep over every system (pgt_ik_sweep): one launch per stage
        # for all 1000 frames, the tile gather at each engine's tuned cell
        # reach; run once eagerly, then captured (about ten graph nodes)
        outs = [(torch.empty_like(j["ov"]), torch.empty_like(j["ok"])) for j in jobs]
        if args.sweep_reach:
            for j in jobs:
                j["e"].set_tuning(cell_reach=args.sweep_reach)
        sw = engine.Sweep([dict(engine=j["e"], nframes=FR, n=j["sy"]["N"], pos=j["sy"]["pos"], diam=j["sy"]["diam"],
                                mass=j["sy"]["mass"], vel=j["sy"]["vel"], virial=o[0], kinetic=o[1])
                           for j, o in zip(jobs, outs)], fields=("virial", "kinetic"), stream_handle=s.cuda_stream)
        sw.run()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            sw.run()
        t_sweep_eager = (time.perf_counter() - t0) / args.steps
        sweep_stats = sw.stats()
        # agreement with the engines' own results (the tile gather sums in the
        # same order as the sweep; the lane gather, where an engine chose it,
        # in another)
        dmax = 0.0
        for j, o in zip(jobs, outs):
            for a, b in ((o[0], j["ov"]), (o[1], j["ok"])):
                dmax = max(dmax, float((a - b).abs().max() / b.abs().max()))
        gs = torch.cuda.CUDAGraph(keep_graph=True)
        with torch.cuda.graph(gs, stream=s):
            sw.enqueue()
        nodes = gs_nodes(gs)
        gs.instantiate()
        sw.finish()
        gs.replay()
        torch.cuda.synchronize()
        sw.finish()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            gs.replay()
        torch.cuda.synchronize()
        t_sweep = (time.perf_counter() - t0) / args.steps
        sw.finish()
        sw.release()
        nframes = FR * len(jobs)
        results[prec] = {"frames_per_s_graph": nframes / t_graph, "frames_per_s_eager": nframes / t_eager,
                         "frames_per_s_graph_%d_streams" % K: nframes / t_graph_k,
                         "frames_per_s_sweep_graph": nframes / t_sweep,
                         "frames_per_s_sweep_eager": nframes / t_sweep_eager,
                         "sweep_graph_nodes": nodes,
                         "sweep_gather_ms_eager": sweep_stats["gather_ms"],
                         "sweep_gather_flops": FLOP_U_2D * sweep_stats["n_contrib"] + FLOP_KIN_2D * sweep_stats["n_ball"],
                         "sweep_total_ms_eager": sweep_stats["total_ms"],
                         "sweep_max_rel_diff_vs_engines": dmax,
                         "contributions_per_s_graph": nc / args.steps / t_graph,
                         "tuning": {str(j["sy"]["N"]): j["tuning"] for j in jobs}}
        log("%s sweep: per-engine graph %.1f frames/s (%d streams: %.1f), eager %.1f; one sweep: graph %.1f "
            "(%s nodes), eager %.1f frames/s, max rel diff vs engines %.2e" %
            (prec, nframes / t_graph, K, nframes / t_graph_k, nframes / t_eager, nframes / t_sweep, nodes,
             nframes / t_sweep_eager, dmax))
        del jobs, g, g2, gs, sw, outs
        torch.cuda.empty_cache()
    best = results["fp64"]["frames_per_s_sweep_graph"]
    line = {"metric": METRIC, "value": round(best, 3), "unit": "frames/s",
            "n_gpus": 1, "steps": args.steps, "warmup": 1,
            "ms_per_step": round(FR * len(C5_N) / best * 1e3, 4),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (seeded lattice+jitter polydisperse glasses, 8 systems x 125 frames resident in HBM)",
            "config": {"workload": "2D polydisperse12, 8 systems N in %s x 125 frames, 128x128 grid each, Octic "
                                   "r_cg=2.0, GL5, virial + kinetic; one sweep (one launch per stage over all "
                                   "systems' frames) captured in one HIP graph" % (list(C5_N),),
                       "frames_per_step": FR * len(C5_N), "grid": "128x128", "parallelism": "single GPU"},
            "roofline": sweep_roofline(results["fp64"]), "cpu_baseline": None,
            "extra": {"fp64": results["fp64"], "fp32_form": results["fp32"],
                      "fp32_over_fp64_graph": results["fp32"]["frames_per_s_graph"] / results["fp64"]["frames_per_s_graph"]}}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", type=int, default=2, choices=[2, 3, 4, 5],
                    help="2: FP64 plain form, 2D N=65536 x 100 frames (default); 3: FP32 precision form, 3D; "
                         "4: learned bond-integral surrogate on the 2D N=2000 workload; "
                         "5: batched sweep over 8 systems of uneven N in one HIP graph")
    ap.add_argument("--c3-sample", type=int, default=2048, help="grid points of the config-3 oracle check")
    ap.add_argument("--c4-epochs", type=int, default=60, help="training epochs of the config-4 surrogate")
    ap.add_argument("--surrogate-epochs", type=int, default=8,
                    help="training epochs of the surrogate comparison in the config-2 line's extra (0: skip it)")
    ap.add_argument("--sweep-streams", type=int, default=4, choices=[1, 2, 4, 8],
                    help="streams the config-5 graph forks its systems onto")
    ap.add_argument("--sweep-reach", type=int, default=0, choices=[0, 1, 2],
                    help="cell reach of every system in the config-5 sweep (0: each engine's tuned reach)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--frames", type=int, default=100)
    ap.add_argument("--n", type=int, default=65536)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the parity sample")
    ap.add_argument("--no-host", action="store_true", help="skip the host-buffer API timing (for profiling runs)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--block", type=int, default=0, help="gather threads per workgroup (0: autotune)")
    ap.add_argument("--fpl", type=int, default=0, help="frames per launch (0: autotune)")
    ap.add_argument("--pl", type=int, default=0, help="grid points per gather lane (0: autotune)")
    ap.add_argument("--reach", type=int, default=1, help="cell reach with --block/--fpl/--pl fixed (1 or 2)")
    ap.add_argument("--gather", default="auto", choices=["auto", "tile", "lane"],
                    help="field gather for the Octic 5-node rule: item-balanced 64-point tiles, one lane per point "
                         "(or group of points), or the faster of the two as the autotuner finds (auto; tile with "
                         "--block/--pl fixed)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 3 device copies of MIB MiB (reference dispatches for HBM byte counters) and exit")
    args = ap.parse_args()

    if args.copy_calibration:
        import torch

        nbytes = args.copy_calibration << 20
        src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": nbytes, "copies": 3}), flush=True)
        return

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (--gpus %d requested)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    if args.config == 3:
        return precision_form(args)
    if args.config == 4:
        return learned_surrogate(args)
    if args.config == 5:
        return batched_sweep(args)

    import torch

    from pyglasstools_amd import engine, workload
    from pyglasstools_amd.autotune import Autotuner

    if engine.device_count() < 1 or not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    dev = torch.device("cuda:0")
    F, N = args.frames, args.n
    log("generating %d frames of N=%d" % (F, N))
    systems = make_frames(F, N)
    L = systems[0]["L"]
    grid = workload.cartesian_grid_2d(L[0], args.grid)
    G = len(grid)
    d0 = float(systems[0]["diam"].max())
    max_rcut = max(1.25 * d0, d0)  # Calculator.h:41-45, from the first frame

    pos = torch.tensor(np.stack([s["pos"] for s in systems]), dtype=torch.float64, device=dev)
    diam = torch.tensor(np.stack([s["diam"] for s in systems]), dtype=torch.float64, device=dev)
    mass = torch.tensor(np.stack([s["mass"] for s in systems]), dtype=torch.float64, device=dev)
    vel = torch.tensor(np.stack([s["vel"] for s in systems]), dtype=torch.float64, device=dev)
    out_v = torch.empty((F, G, 2, 2), dtype=torch.float64, device=dev)
    out_k = torch.empty((F, G, 2, 2), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    eng = engine.Engine(dim=2, max_rcut=max_rcut, precision=args.precision)
    eng.set_box(L, [1, 1, 0])
    eng.set_grid(grid)

    def step():
        eng.compute_frames(F, N, pos, diam, mass, vel, None, fields=("virial", "kinetic"), d_virial=out_v,
                           d_kinetic=out_k)

    tuned, trials = None, []
    if args.block and args.fpl and args.pl:
        mode = "tile" if args.gather == "auto" else args.gather
        eng.set_gather_mode(mode)
        eng.set_tuning(args.block, args.fpl, args.pl, args.reach)
        tuned = {"gather": mode, "threads_per_block": args.block, "frames_per_launch": args.fpl,
                 "points_per_lane": args.pl, "cell_reach": args.reach}
    else:
        log("autotuning gather kind x cell reach x points per lane x threads per workgroup x frames per launch")
        gathers = ("lane", "tile") if args.gather == "auto" else (args.gather,)
        tuned, trials = Autotuner().tune(eng, step, ("c2", N, G, F, args.precision, args.gather), F,
                                         sync=torch.cuda.synchronize, gathers=gathers)
        log("tuned: %s" % tuned)
    mode = tuned["gather"]

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # (a) one compute_frames per step, the host waiting for each job (its
    # counters read back, the bond capacity checked); the jobs' HIP events
    # give the gather's launch time for the roofline
    gather_ms, n_contrib, n_ball, launches = 0.0, 0, 0, 0
    fpl = tuned["frames_per_launch"]
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        st = eng.stats()
        gather_ms += st["gather_ms"]
        n_contrib += st["n_contrib"]
        n_ball += st["n_ball"]
        launches += (F + fpl - 1) // fpl
    torch.cuda.synchronize()
    elapsed_sync = time.perf_counter() - t0
    # (b) the value: the same jobs enqueued back to back on one stream
    # (pgt_ik_enqueue_frames: no host synchronisation between jobs), one
    # synchronisation at the end, then pgt_ik_finish checks every job's bond
    # capacity and returns the counts, which must equal (a)'s
    s = torch.cuda.Stream()
    eng.set_stream(s.cuda_stream)
    with torch.cuda.stream(s):
        step()
    torch.cuda.synchronize()
    eng.finish()
    t0 = time.perf_counter()
    with torch.cuda.stream(s):
        for _ in range(args.steps):
            eng.enqueue_frames(F, N, pos, diam, mass, vel, None, fields=("virial", "kinetic"), d_virial=out_v,
                               d_kinetic=out_k)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    eng.finish()
    stq = eng.stats()
    if stq["n_contrib"] != n_contrib or stq["n_ball"] != n_ball:
        raise RuntimeError("pipelined jobs counted %d / %d contributions / ball pairs, the synchronous ones %d / %d"
                           % (stq["n_contrib"], stq["n_ball"], n_contrib, n_ball))
    eng.release_enqueued()
    eng.set_stream(0)
    fps = F * args.steps / elapsed
    log("timed %d steps in %.3f s: %.1f frames/s pipelined (%.1f with a host wait per job)" %
        (args.steps, elapsed, fps, F * args.steps / elapsed_sync))

    flops = FLOP_U_2D * n_contrib + FLOP_KIN_2D * n_ball
    avg_launch_s = gather_ms / 1e3 / launches
    achieved = flops / launches / avg_launch_s / 1e12
    traffic, traffic_fpl = traffic_from_profile(mode)
    if traffic is not None and traffic_fpl != fpl:
        traffic = traffic * fpl / traffic_fpl  # per-launch bytes scale with frames per launch
    kname = "k_gather_tile (item-balanced tile field gather)" if mode == "tile" else "k_gather (lane field gather)"
    roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": traffic,
                "kernel": kname, "compute": "FP64 vector ALU (published MI355X FP64 peak 78.6 TFLOP/s)",
                "avg_launch_ms": round(avg_launch_s * 1e3, 4), "flops_per_launch": flops / launches}
    if traffic is not None:
        # HBM bytes of the gather (PMC counters) over its launch time, against 8 TB/s
        roofline["hbm_gbs"] = round(traffic / avg_launch_s / 1e9, 1)
        roofline["hbm_frac"] = round(traffic / avg_launch_s / 1e9 / PEAK_HBM_GBS, 4)
        roofline["algorithmic_bytes_per_launch"] = fpl * (N * (16 + 8 + 8 + 24) + 2 * G * 32)

    extra = {"frames": F, "n_particles": N, "grid_points": G, "contributions_per_s": n_contrib / elapsed,
             "contributions_per_frame": n_contrib / (F * args.steps), "kinetic_pairs_per_frame": n_ball / (F * args.steps),
             "gather_share_of_step": gather_ms / 1e3 / elapsed,
             "frames_per_s_host_wait_per_job": F * args.steps / elapsed_sync,
             "timing": "value: %d jobs of %d frames enqueued back to back on one stream, one synchronisation "
                       "(pgt_ik_enqueue_frames, then pgt_ik_finish checks the bond capacity and counts); "
                       "frames_per_s_host_wait_per_job: one pgt_ik_compute_frames per job" % (args.steps, F),
             "tuning": tuned, "precision": args.precision}
    if trials:
        extra["autotune_trials"] = trials

    # host-buffer API (what IrvingKirkwood.compute uses): per-frame upload,
    # compute and download through pgt_ik_compute, PCIe copies included
    if not args.no_host:
        host = engine.Engine(dim=2, max_rcut=max_rcut, precision=args.precision)
        host.set_box(L, [1, 1, 0])
        host.set_grid(grid)
        nh = min(F, 10)
        for f in range(2):
            host.set_particles(systems[f]["pos"], systems[f]["diam"], systems[f]["mass"], systems[f]["vel"])
            host.compute(fields=("virial", "kinetic"))
        t0 = time.perf_counter()
        for f in range(nh):
            host.set_particles(systems[f]["pos"], systems[f]["diam"], systems[f]["mass"], systems[f]["vel"])
            host.compute(fields=("virial", "kinetic"))
        extra["host_api_fps_pcie_inclusive"] = nh / (time.perf_counter() - t0)
        host.close()

    if not args.no_host:
        extra["driver_script"] = driver_script_time()
        log("driver script: %s" % extra["driver_script"])

    # the metric's other half: the learned surrogate against plain quadrature
    # on the config-1 workload, with a short training schedule (config 4 runs
    # the full one)
    if args.surrogate_epochs > 0:
        extra["surrogate_vs_plain_c1"], _, _ = surrogate_comparison(args.surrogate_epochs, 50)

    cpu = None
    if not args.no_cpu:
        step()
        torch.cuda.synchronize()
        extra["parity_frame0"] = parity_sample(systems, grid, max_rcut, out_v[0].cpu().numpy(), out_k[0].cpu().numpy())
        log("parity sample: %s" % extra["parity_frame0"])
        res, threads, hinfo = cpu_baseline(systems, grid, max_rcut, args.cpu_seconds)
        log("cpu oracle: %s on %s" % (res, hinfo))
        cpu = {"value": round(res["clean"]["fps"], 4), "unit": "frames/s", "cores": threads, "kind": "port",
               "sample": "%d of the %d frames after one warm-up frame, full 256x256 grid, median per-frame "
                         "time; FP64 oracle (C restatement of IrvingKirkwood::compute, OpenMP over grid points, "
                         "no per-triple allocations)" % (res["clean"]["frames"], F),
               "cpu_model": hinfo["model"], "cpus_online": hinfo["cpus_online"],
               "cpus_usable": hinfo["cpus_usable"], "build": hinfo["build"]}
        extra["cpu_faithful_fps"] = res["faithful"]["fps"]
        extra["cpu_faithful_sample_frames"] = res["faithful"]["frames"]
        extra["speedup_vs_cpu_clean"] = fps / res["clean"]["fps"]
        extra["speedup_vs_cpu_faithful"] = fps / res["faithful"]["fps"]

    line = {"metric": METRIC, "value": round(fps, 3), "unit": "frames/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 4), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64" if args.precision == "fp64" else "f32+f64",
            "data": "synthetic (seeded lattice+jitter polydisperse glass, 100 frames resident in HBM)",
            "config": {"workload": "2D polydisperse12 N=65536, 100 frames, 256x256 grid, Octic r_cg=2.0, GL5, "
                                   "FP64 virial + kinetic stress", "frames_per_step": F, "n_particles": N,
                       "grid": "%dx%d" % (args.grid, args.grid), "parallelism": "single GPU"},
            "roofline": roofline, "cpu_baseline": cpu, "extra": extra}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
