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
  cpu_baseline   the C oracle (a restatement of the serial Fortran) on one
                 host core over a bounded sample of the same workload
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "nemolite2d-cuda_amd"))

METRIC = "cell-updates/s at 4096² FP64 vs FP32/ensemble/closure; % HBM peak; SSH RMS err"
HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL = 80.0


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


def load_traffic():
    p = os.path.join(ROOT, "profiles", "traffic_step_fused_4096.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f).get("bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=4096, help="interior grid size (jpi = jpj)")
    ap.add_argument("--cpu-steps", type=int, default=8, help="CPU baseline sample (0: skip)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also reset the state 3 more times: each reset issues device-to-"
                         "device copies of the five padded state arrays (plain copy "
                         "kernels of known size) whose HBM counters calibrate the "
                         "FETCH_SIZE/WRITE_SIZE readings of the step kernel")
    ap.add_argument("--variant", default="auto",
                    help="fused-step kernel (nemo2d.VARIANTS); results do not depend on it")
    ap.add_argument("--sweep", action="store_true",
                    help="time every fused-step variant (200 steps each) and report them "
                         "under 'variants'")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (--gpus %d requested); the "
              "multi-GPU decomposition is a design, not an implementation" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import nemo2d
    n, K, W = args.n, args.steps, args.warmup
    m = nemo2d.Model(n, n, dx=1000.0, dy=1000.0, dep_const=100.0, nit000=1, nitend=K,
                     irecord=K + 1, rdt=20.0, cbfr=0.0015, visc=0.1)
    m.set_variant(args.variant)
    rest = nemo2d.initial_state(nemo2d.default_grid(n, n))

    # warm-up, then reset to rest
    m.record_checksums(max(K, W))
    for istp in range(1, W + 1):
        m.step(istp)
    m.checksums()
    m.set_state(rest)
    m.record_checksums(K)

    # timed region: K steps with per-step checksums
    t0 = time.perf_counter()
    m.timer_start()
    m.run(1, K)
    ms = m.timer_stop()
    wall = time.perf_counter() - t0
    log = m.checksum_log(K)
    final = m.checksums()

    # kernel region: the same K steps without checksum recording, so the
    # region holds only fused-step launches (its average is the kernel's)
    m.set_state(rest)
    m.record_checksums(0)
    m.timer_start()
    m.run(1, K)
    ms_kernel = m.timer_stop() / K

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
    traffic = load_traffic()
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
                     "kernel": "k_step_march", "kernel_ms": ms_kernel,
                     "bytes_per_launch": BYTES_PER_CELL * cells},
    }
    if variants:
        out["variants_ms_per_step"] = variants
    if args.cpu_steps > 0:
        out["cpu_baseline"] = cpu_baseline(n, args.cpu_steps)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
