#!/usr/bin/env python3
"""Benchmark: NEMOLite2D FP64 drop-in at 4096x4096 on one MI355X.

A step is one time step of the model (continuity, momentum, boundary
conditions, next) over the whole 4096^2 T-grid, through the library's fused
step, with the per-step sum|ssh|, sum|u|, sum|v| checksums recorded on the
device.  The workload is the 4096^2, 1000-step run from rest (jpiglo =
jpjglo = 4098, dx = dy = 1000 m, depth 100 m, rdt = 20 s, cbfr = 0.0015,
visc = 0.1): after W warm-up steps the state is reset to rest and steps
1..K are timed with HIP events on the model's stream (K = 1000 by default,
the whole run).

Prints one JSON line (metric/unit as the benchmark definition names them):
  value          cell-updates/s = 4096*4096*K / t
  roofline       the fused step kernel against MI355X HBM (8 TB/s), with
                 80 algorithmic bytes per cell-update (5 FP64 state arrays
                 read once and written once) over its average launch time
  cpu_baseline   the reference's serial Fortran model itself (built from its
                 sources into oracle/_ref with the reference's own -Ofast
                 optimisation; the IEEE -O2 build beside it) on one host
                 core over a bounded sample of the same workload (steps
                 1..16); the C oracle (a restatement of it) where neither
                 binary is there
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "nemolite2d-cuda_amd"))

METRIC = "cell-updates/s at 4096² FP64 vs FP32/ensemble/closure; % HBM peak; SSH RMS err"
HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL = 80.0


REF_EXE = os.path.join(ROOT, "oracle", "_ref", "nemolite2d_ref")
REF_EXE_FAST = os.path.join(ROOT, "oracle", "_ref", "nemolite2d_ref_fast")
REF_FLAGS = {REF_EXE: "flang -O2 -ffp-contract=off: IEEE double, subnormals kept (no FTZ/DAZ), no FMA "
                      "contraction",
             REF_EXE_FAST: "flang -O3 -ffast-math -march=x86-64-v3 (the reference's own -Ofast): "
                           "flush-to-zero, FMA contraction, reassociation"}
REF_NAMELIST = """&namctl
   jpiglo = {g}
   jpjglo = {g}
   jphgr_msh = 1
   dx = 1000.0
   dy = 1000.0
   dep_const = 100.0
   nit000 = 1
   nitend = {n}
   irecord = {r}
   rdt = 20.0
   cbfr = 0.0015
   visc = 0.1
/
"""


def progress(msg):
    """A progress line on stderr (the JSON line alone goes to stdout)."""
    print("bench.py: " + msg, file=sys.stderr, flush=True)


def host_facts():
    """The host the CPU baseline ran on: CPU model, logical CPUs and the CPUs
    this process may use (the timed binary's floating-point mode is in its
    "flags")."""
    model = None
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
    except (AttributeError, OSError):
        usable = None
    return {"cpu_model": model, "logical_cpus": os.cpu_count(), "usable_cpus": usable}


def cpu_baseline_reference(n, steps, exe=REF_EXE):
    """The reference's own serial model (a binary oracle/_ref builds from
    original/nemolite2d.f90) on one core: steps 1..steps of the same run from
    rest, timed by its own dl_timer 'Time-stepping' region
    (original/nemolite2d.f90:70-79).  Its unconditional istp=0 snapshot
    (:67-68) goes to /dev/null.  None if the binary is absent or fails."""
    import shutil
    import subprocess
    import tempfile
    if not os.access(exe, os.X_OK):
        return None
    d = tempfile.mkdtemp(prefix="nemo2d_ref_")
    try:
        with open(os.path.join(d, "namelist"), "w") as f:
            f.write(REF_NAMELIST.format(g=n + 2, n=steps, r=steps + 1))
        os.symlink(os.devnull, os.path.join(d, "go2d_00000.dat"))
        core = sorted(os.sched_getaffinity(0))[0]
        env = dict(os.environ, OMP_NUM_THREADS="1")
        p = subprocess.run([exe], cwd=d, env=env, capture_output=True, text=True, timeout=600,
                           preexec_fn=lambda: os.sched_setaffinity(0, {core}))
        if p.returncode != 0:
            print("bench.py: reference binary failed (%d): %s" % (p.returncode, p.stderr[-500:]),
                  file=sys.stderr)
            return None
        rows = [l.split() for l in p.stdout.splitlines() if l.startswith("Time-stepping")]
        t = float(rows[-1][2])
    except (OSError, subprocess.SubprocessError, IndexError, ValueError) as e:
        print("bench.py: reference binary not timed: %r" % (e,), file=sys.stderr)
        return None
    finally:
        shutil.rmtree(d, ignore_errors=True)
    return {"value": n * n * steps / t, "unit": "cell-updates/s", "cores": 1, "kind": "reference",
            "sample": "the reference's serial model (original/nemolite2d.f90 and the dl_esm_inf/"
                      "dl_timer sources it links, built into oracle/_ref by %s), steps 1..%d of "
                      "the %dx%d run from rest, its dl_timer 'Time-stepping' region: %.2f s, one "
                      "core" % (REF_FLAGS[exe].split(":")[0], steps, n, n, t),
            "flags": REF_FLAGS[exe]}


def cpu_baseline(n, steps):
    """C oracle on one core: steps 1..steps of the same run from rest."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    try:
        os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
    except (AttributeError, OSError):
        pass
    from oracle import Oracle
    o = Oracle(n, n)
    t = o.run_timed(1, steps)
    del o
    return {"value": n * n * steps / t, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": "C oracle (restatement of original/nemolite2d.f90, gcc -O2 "
                      "-ffp-contract=off), steps 1..%d of the %dx%d run from rest, %.1f s, "
                      "per-step checksums off" % (steps, n, n, t)}


def form_fp32(n, steps, seed=1):
    """Single-precision form on a seamount basin against the FP64 path on
    the same inputs."""
    import nemo2d
    grid, _ = nemo2d.seamount_grid(n, n, seed)
    rest = nemo2d.initial_state(grid)
    out = {}
    for prec in ("fp64", "fp32"):
        m = nemo2d.Model(n, n, grid=grid, state=rest, precision=prec, rdt=20.0, cbfr=0.0015,
                         visc=0.1)
        m.run(1, 2)
        m.set_state(rest)
        m.timer_start()
        m.run(1, steps)
        ms = m.timer_stop()
        out[prec] = (ms, m.retrieve(), m.diagnostics())
        m.close()
    err = nemo2d.field_errors(out["fp32"][1], out["fp64"][1])
    cells = float(n) * n * steps
    return {"workload": "%dx%d seamount basin (32 seeded Gaussian mounts, depth 100 m minus "
                        "heights U(10,60) m, radii U(20,200) km, floor 10 m), %d steps, "
                        "rdt=20 s, cbfr=0.0015" % (n, n, steps),
            "fp32_cell_updates_per_s": cells / (out["fp32"][0] * 1e-3),
            "fp64_cell_updates_per_s": cells / (out["fp64"][0] * 1e-3),
            "errors_vs_fp64": err,
            "volume": {"fp64": out["fp64"][2][0], "fp32": out["fp32"][2][0]},
            "energy": {"fp64": out["fp64"][2][1], "fp32": out["fp32"][2][1]}}


def form_ensemble(nmem, steps, seed=7, graph=True, tiling=None):
    """Batched basins: one kernel per step, graph-captured step replayed,
    autotuned tiling (or the given one); FP64 and FP32 on the same members."""
    import nemo2d
    mem = nemo2d.ensemble_members(nmem, seed)
    res = {}
    states = {}
    for prec in (64, 32):
        e = nemo2d.Ensemble(mem, rdt=12.0, precision=prec)
        if tiling is None:
            progress("ensemble fp%d: tuning" % prec)
            choice, _ = e.tune(nsteps=10)
        else:
            choice = tiling
            e.set_tiling(choice)
        progress("ensemble fp%d: tiling %d, %d steps%s" % (prec, choice, steps, " (graph)" if graph else ""))
        ms = e.run(1, steps, graph=graph)
        res[prec] = {"cell_updates_per_s": e.cells * steps / (ms * 1e-3), "ms": ms,
                     "tiling": nemo2d.Ensemble.tiling_shape(choice)}
        states[prec] = [e.retrieve(k) for k in range(0, nmem, max(1, nmem // 16))]
        cells = e.cells
        e.close()
    errs = [nemo2d.field_errors(a, b) for a, b in zip(states[32], states[64])]
    rel = {f: float(np.sqrt(np.mean([(x[f]["rms"] / max(x[f]["ref_rms"], 1e-300)) ** 2 for x in errs])))
           for f in ("ssh", "u", "v")}
    return {"workload": "%d square basins, side round(exp U(ln64, ln1024)) (%d cells in all), depth "
                        "U(50,200) m, tide amplitude U(0.1,0.3) m, period U(11,13) h, seed %d, %d "
                        "steps, rdt=12 s" % (nmem, cells, seed, steps),
            "fp64": res[64], "fp32": res[32],
            "fp32_rms_error_relative_to_fp64_rms (16 members)": rel}


# The closure workload: a 1024^2 coarse basin (dx = 2 km, rdt = 40 s) and the
# 4-layer corrector, trained on pairs from 64 seeded 2048^2 seamount fine runs
# (dx = 1 km, rdt = 20 s; the seamounts of the 8192^2 FP32 workload: 32
# mounts, heights U(10,60) m, radii U(20,200) km) restricted to 1024^2, 313
# windows of 10 coarse steps each (20,032 pairs).
CLOSURE_BATHY = dict(n_mounts=32, radius=(20e3, 200e3), height=(10.0, 60.0))
CLOSURE_LR = 1e-3
CLOSURE_ROLLOUT = 2   # the prediction model restarts from the fine state every 2nd window
CONV_FLOP_PER_PIXEL = 2 * 9 * 32 * 32   # one 32 -> 32 layer, forward or weight gradient
MFMA_BF16_PEAK = 2.5e15


def conv_kernel_rates(net, nc, reps=20):
    """The closure network's convolution kernels of one 32 -> 32 layer at the
    training batch (4 images nc x nc): time per launch, MFMA flop/s and HBM
    bytes/s against the MI355X peaks."""
    rng = np.random.default_rng(0)
    P = 4 * nc * nc
    xp = np.zeros((P, 8), np.float32)
    xp[:, :4] = rng.standard_normal((P, 4))
    xg = net.to_bf16(xp)
    kt = net.time_kernels(xg, 4, nc, nc, reps=reps)
    flop = CONV_FLOP_PER_PIXEL * P
    kernels = {}
    for name, t in kt.items():
        # HBM bytes per pixel: forward reads X (64) and writes z, act (64 + 64)
        # (in inference act only); data gradient reads dZ, z below (64 + 64)
        # and writes dz below (64); weight gradient reads X and dZ (64 + 64)
        bpp = {"conv_fwd": 192, "conv_fwd_inference": 128, "conv_bwd_data_gelu": 192,
               "conv_bwd_weight": 128}[name]
        kernels[name] = {"us": t * 1e6, "tflops": flop / t / 1e12, "frac_of_bf16_mfma_peak": flop / t / MFMA_BF16_PEAK,
                         "hbm_gbs": bpp * P / t / 1e9, "frac_of_hbm": bpp * P / t / 1e9 / HBM_PEAK_GBS}
    return kernels


def form_conv(nc=1024, reps=20):
    """The convolution kernels alone (an untrained network of the closure's
    shape), for kernel traces and counters."""
    from nemo2d.closure import ConvNet
    return {"workload": "closure network 3x3 convolutions, one 32 -> 32 layer, batch 4 x %dx%d (BF16, "
                        "FP32 accumulation), %d launches each" % (nc, nc, reps),
            "conv_kernels_batch_4x%dx%d" % (nc, nc): conv_kernel_rates(ConvNet(seed=0), nc, reps)}


def form_closure(sc, eval_windows=60, train=None):
    """Coarse (1024^2) + learned closure against the fine (2048^2) solver:
    training on the device, the held-out basin judged by the fine run,
    throughput of the three runs, and the convolution kernels' rates at the
    training batch (4 x 1024^2).  train = (basins, windows) overrides the
    training's size (a short run at the full grid size, for kernel traces)."""
    import nemo2d
    from nemo2d import closure_pipeline as cp
    nc = 1024 if sc >= 1 else max(64, int(1024 * sc) // 32 * 32)
    basins = max(4, int(round(64 * min(sc, 1.0))))
    windows = max(10, int(round(313 * min(sc, 1.0))))
    if train:
        basins, windows = train
    t0 = time.perf_counter()
    clos, info = cp.train_online(nc, range(basins), windows, group=4, bathy=CLOSURE_BATHY, lr=CLOSURE_LR,
                                 rollout=CLOSURE_ROLLOUT, log=progress)
    t_train = time.perf_counter() - t0
    held = cp.Setup(nc, 1000, **CLOSURE_BATHY)
    t0 = time.perf_counter()
    ep, ec, fell = cp.evaluate_device(held, clos, eval_windows, tol=1.0, judge_every=1)
    _, ec_free, _ = cp.evaluate_device(held, clos, eval_windows, tol=float("inf"))
    t_eval = time.perf_counter() - t0
    first = [k for k in range(len(ec_free)) if not ec_free[k] < ep[k]]
    # throughput: plain coarse, coarse with the on-device correction every 10
    # steps, fine; each over the evaluation's windows
    K = cp.K_STEPS
    rates = {}
    for name in ("coarse", "coarse+closure", "fine"):
        m = held.fine_model() if name == "fine" else held.coarse_model()
        nst = (2 if name == "fine" else 1) * K * eval_windows
        m.run(1, 2)
        m.checksums()
        t1 = time.perf_counter()
        if name == "coarse+closure":
            for w in range(eval_windows):
                m.run(K * w + 1, K * w + K)
                clos.correct_model(m)
        else:
            m.run(1, nst)
        m.checksums()
        dt = time.perf_counter() - t1
        n = held.nf if name == "fine" else held.nc
        rates[name] = n * n * nst / dt
        m.close()
    kernels = conv_kernel_rates(clos.net, nc)
    return {"workload": "coarse %dx%d (dx=2 km, rdt=40 s) + 4-layer 3x3-conv closure (4->32->32->32->3, "
                        "GELU, BF16 MFMA 32x32x16, FP32 accumulation) applied every %d steps; trained on "
                        "the device on %d pairs from %d seeded seamount fine runs %dx%d (dx=1 km, rdt=20 s, "
                        "32 mounts, heights 10-60 m, radii 20-200 km) restricted to %dx%d, %d windows "
                        "each, batch 4, one Adam step per batch, the prediction restarted from the fine "
                        "state every %d windows and corrected by the network in between; judged on a "
                        "held-out basin (seed 1000) "
                        "by the fine run for %d windows" % (nc, nc, K, info["pairs"], basins, 2 * nc, 2 * nc,
                                                            nc, nc, windows, CLOSURE_ROLLOUT, eval_windows),
            "ssh_rms_vs_fine": {"coarse": float(ep.mean()), "coarse+closure (judged every window)": float(ec.mean()),
                                "coarse+closure (no judge)": float(np.nanmean(ec_free)) if np.isfinite(ec_free).any() else None},
            "closure_better_than_coarse_until_window": first[0] if first else eval_windows,
            "fallback_fired": bool(fell),
            "train_loss_first_last": [info["losses"][0][1], info["losses"][-1][1]],
            "cell_updates_per_s": rates,
            "fine_equivalent_rate_of_coarse+closure": rates["coarse+closure"] * 8.0,
            "conv_kernels_batch_4x%dx%d" % (nc, nc): kernels,
            "seconds": {"train_total": t_train, "eval": t_eval, **info["seconds"]}}


# the fused-step kernel each variant runs on the uniform benchmark grid
# (march12_rc: <3> divides by reciprocals with one correction, as on the
# benchmark grid, whose divisors all allow it; <2> with two elsewhere)
KERNELS = {"march12_rc": "k_step_march_rc<3>", "march8_rc": "k_step_march_rc<3>",
           "march12_dpfb": "k_step_march_dp<12, true, true>",
           "march12_dpf": "k_step_march_dp<12, true, false>", "march12_dpb": "k_step_march_dp<12, false, true>",
           "march12_dp": "k_step_march_dp<12, false, false>", "march12_shn": "k_step_march_sh<12>",
           "march16": "k_step_march", "cell": "k_step_fused<nemo2d::StUniform, false>",
           "cell_flin": "k_step_fused<nemo2d::StUniform, true>"}


def load_traffic(v, n):
    """Measured HBM bytes per launch of this variant's kernel at this size
    (profiles/traffic_step_<variant>_<n>.json), else None."""
    p = os.path.join(ROOT, "profiles", "traffic_step_%s_%d.json" % (v, n))
    if not os.path.exists(p):
        return None
    with open(p) as f:
        d = json.load(f)
    return d.get("bytes_per_launch") if d.get("variant") == v else None


def run_forms(sc, which=("fp32_mixed", "ensemble", "closure"), graph=True, keep_going=False,
              closure_train=None, ens_tiling=None, ens_steps=None):
    """The forms' figures; with keep_going a form that raises is reported as
    {"error": ...} (printed to stderr too) and the others still run."""
    runs = {"fp32_mixed": lambda: form_fp32(int(8192 * sc) // 64 * 64 or 64, max(10, int(5000 * sc))),
            "conv": lambda: form_conv(),
            "ensemble": lambda: form_ensemble(max(8, int(256 * sc)), ens_steps or max(10, int(3000 * sc)),
                                              graph=graph, tiling=ens_tiling),
            "closure": lambda: form_closure(sc, train=closure_train)}
    out = {}
    for name in ("fp32_mixed", "ensemble", "closure", "conv"):
        if name not in which:
            continue
        progress("form %s" % name)
        if not keep_going:
            out[name] = runs[name]()
            continue
        try:
            out[name] = runs[name]()
        except Exception as e:  # the headline figure is still reported
            print("bench.py: form %s failed: %r" % (name, e), file=sys.stderr, flush=True)
            out[name] = {"error": repr(e)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=4096, help="interior grid size (jpi = jpj)")
    ap.add_argument("--cpu-steps", type=int, default=16, help="CPU baseline sample (0: skip)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also reset the state 3 more times: each reset issues device-to-"
                         "device copies of the five padded state arrays (plain copy "
                         "kernels of known size) whose HBM counters calibrate the "
                         "FETCH_SIZE/WRITE_SIZE readings of the step kernel")
    ap.add_argument("--variant", default="auto",
                    help="fused-step kernel (nemo2d.VARIANTS); results do not depend on it")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the single-precision, ensemble and closure forms")
    ap.add_argument("--forms-scale", type=float, default=1.0,
                    help="scale of the forms' workloads (1.0: 8192^2 x 5000 FP32, 256-basin "
                         "ensemble x 3000, closure 1024^2/2048^2 with 64 basins x 313 windows)")
    ap.add_argument("--form", choices=("fp32_mixed", "ensemble", "closure", "conv"),
                    help="run only this form (at --forms-scale) and print its JSON line; "
                         "for per-form kernel traces (conv: the closure's convolution "
                         "kernels alone)")
    ap.add_argument("--closure-train", type=lambda v: tuple(int(x) for x in v.split(",")),
                    help="with --form closure: BASINS,WINDOWS of training instead of the "
                         "workload's 64,313 (grid sizes unchanged; for kernel traces)")
    ap.add_argument("--ens-tiling", type=int,
                    help="with --form ensemble: run this tiling instead of the autotuner's pick")
    ap.add_argument("--ens-steps", type=int,
                    help="with --form ensemble: steps per run instead of the workload's 3000 "
                         "(for counter passes)")
    ap.add_argument("--no-graph", action="store_true",
                    help="with --form ensemble: launch the steps one by one instead of "
                         "replaying the captured graph")
    ap.add_argument("--sweep", action="store_true",
                    help="time every fused-step variant (200 steps each) and report them "
                         "under 'variants'")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (--gpus %d requested); the "
              "multi-GPU decomposition is a design, not an implementation" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    if args.form:
        print(json.dumps({args.form: run_forms(args.forms_scale, (args.form,), graph=not args.no_graph,
                                               closure_train=args.closure_train,
                                               ens_tiling=args.ens_tiling,
                                               ens_steps=args.ens_steps)[args.form]}),
              flush=True)
        return

    import nemo2d
    n, K, W = args.n, args.steps, args.warmup
    m = nemo2d.Model(n, n, dx=1000.0, dy=1000.0, dep_const=100.0, nit000=1, nitend=K,
                     irecord=K + 1, rdt=20.0, cbfr=0.0015, visc=0.1)
    m.set_variant(args.variant)
    rest = nemo2d.initial_state(nemo2d.default_grid(n, n))

    # warm-up: steps, the recorded log's read, the on-demand checksums and a
    # read-back of the state (the first launch of each kernel loads its code,
    # and the process's first large device-to-host copy sets up its copy path,
    # ~6.5 ms), then reset to rest
    m.record_checksums(max(K, W))
    for istp in range(1, W + 1):
        m.step(istp)
    m.checksum_log(W)
    m.checksums()
    m.retrieve()
    m.set_state(rest)
    m.record_checksums(K)

    # timed region: K steps with per-step checksums, including reading the
    # checksum log back (the step kernels leave block partials per step; the
    # log read reduces them all)
    t0 = time.perf_counter()
    m.timer_start()
    m.run(1, K)
    log = m.checksum_log(K)
    ms = m.timer_stop()
    wall = time.perf_counter() - t0
    final = m.checksums()

    # kernel region: the same K steps without checksum recording, so the
    # region holds only fused-step launches (its average is the kernel's)
    m.set_state(rest)
    m.record_checksums(0)
    m.timer_start()
    m.run(1, K)
    ms_kernel = m.timer_stop() / K
    ran = m.last_variant()

    variants = None
    if args.sweep:
        variants = {}
        for name in nemo2d.VARIANTS:
            if name == "auto":
                continue
            m.set_variant(name)
            m.set_state(rest)
            m.run(1, 5)
            m.set_state(rest)
            m.timer_start()
            m.run(1, 200)
            variants[name] = m.timer_stop() / 200
        m.set_variant(args.variant)
    if args.copy_calibration:
        for _ in range(3):
            m.set_state(rest)
    m.close()

    cells = float(n) * n
    value = cells * K / (ms * 1e-3)
    achieved = BYTES_PER_CELL * cells / (ms_kernel * 1e-3) / 1e9
    traffic = load_traffic(ran, n)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": ms / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: the model's own rest state and tidal forcing (no input data)",
        "config": {"workload": "nemolite2d fp64 drop-in, %dx%d T-grid (jpiglo=jpjglo=%d), "
                               "%d steps from rest, rdt=20 s, cbfr=0.0015, visc=0.1, "
                               "per-step checksums on" % (n, n, n + 2, K),
                   "jpi": n, "jpj": n, "path": "fused step (nemo2d_step)"},
        "wall_s": wall,
        "final_checksums": {"ssh": float(final[0]), "u": float(final[1]), "v": float(final[2])},
        "checksum_last_step": [float(x) for x in log[-1]] if len(log) else None,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "achieved_basis": "algorithmic: 80 B per cell-update (five FP64 state arrays "
                                       "read and written once) over the kernel's mean launch time, "
                                       "HIP events on the model's stream, %d launches" % K,
                     "traffic_source": ("profiles/traffic_step_%s_%d.json: rocprofv3 FETCH_SIZE and "
                                        "WRITE_SIZE of this kernel at this size (a committed "
                                        "measurement, not this run)" % (ran, n)) if traffic else None,
                     "kernel": KERNELS.get(ran, "fused step, variant %s" % ran), "variant": ran,
                     "kernel_ms": ms_kernel,
                     "bytes_per_launch": BYTES_PER_CELL * cells},
    }
    # practical ceiling: plain streaming of the same five-in/five-out bytes
    # over arrays of the padded size ((jpj+4) rows of the 16-padded pitch)
    pitch = (16 + n + 3 + 15) // 16 * 16
    out["stream_5r5w_gbs"] = nemo2d.stream_bandwidth((n + 4) * pitch)
    out["roofline"]["frac_of_stream"] = achieved / out["stream_5r5w_gbs"]
    if variants:
        out["variants_ms_per_step"] = variants
    if not args.no_forms:
        out["forms"] = run_forms(args.forms_scale, keep_going=True)
    if args.cpu_steps > 0:
        progress("CPU baseline")
        # the reference binary built with the reference's own optimisation
        # (-Ofast), beside the IEEE build the parity fixtures come from; the
        # oracle port where neither is there
        fast = cpu_baseline_reference(n, args.cpu_steps, REF_EXE_FAST)
        ieee = cpu_baseline_reference(n, args.cpu_steps, REF_EXE)
        out["cpu_baseline"] = fast or ieee or cpu_baseline(n, args.cpu_steps)
        if fast and ieee:
            out["cpu_baseline"]["ieee_build"] = {k: ieee[k] for k in ("value", "sample", "flags")}
        out["cpu_baseline"]["host"] = host_facts()
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
