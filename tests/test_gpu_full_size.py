"""The forms at the sizes of their benchmark workloads (one GPU):

* single precision on the 8192^2 seamount basin for 5000 steps, against the
  FP64 path on the same inputs, with the FP32-ssh ablation;
* the 256-basin ensemble (sides 64..1024) for 3000 steps in FP64 and FP32
  with the autotuned tiling: sampled FP64 members bitwise equal to their
  single runs, FP32 members within 1e-4 of the FP64 RMS of each field
  (measured about 1e-6 over 16 members).
"""
import numpy as np
import pytest

import nemo2d

pytestmark = pytest.mark.gpu

N3, STEPS3 = 8192, 5000
NMEM, STEPS4, RDT4 = 256, 3000, 12.0


def run_state(grid, n, nsteps, precision, compensated=True):
    print("config-3 run: %dx%d %s compensated=%s %d steps" % (n, n, precision, compensated, nsteps),
          flush=True)
    with nemo2d.Model(n, n, grid=grid, precision=precision, compensated=compensated, rdt=20.0,
                      cbfr=0.0015, visc=0.1) as m:
        m.run(1, nsteps)
        r = m.retrieve()
        d = m.diagnostics()
    return {k: r[k] for k in ("sshn", "un", "vn")}, d


def test_fp32_8192_seamount_5000_steps_against_fp64():
    """Bounds: FP32 (FP64 ssh accumulation) ssh/u/v RMS error below 1e-4 of
    the FP64 field's RMS (measured 4e-7 for ssh, 2e-6 for u, 9e-7 for v); the
    FP32-ssh ablation errs more in ssh; the FP32 volume sum(e12t*ssh) within
    1e-5 relative of the FP64 volume (measured 6e-7)."""
    grid, _ = nemo2d.seamount_grid(N3, N3, seed=1)
    ref, d64 = run_state(grid, N3, STEPS3, "fp64")
    f32, d32 = run_state(grid, N3, STEPS3, "fp32")
    e = nemo2d.field_errors(f32, ref)
    del f32
    raw, draw = run_state(grid, N3, STEPS3, "fp32", compensated=False)
    e_raw = nemo2d.field_errors(raw, ref)
    print("fp32 vs fp64:", e)
    print("fp32 (FP32 ssh) vs fp64:", e_raw)
    print("volume fp64 %.17e fp32 %.17e fp32-ssh %.17e" % (d64[0], d32[0], draw[0]))
    for k in ("ssh", "u", "v"):
        assert e[k]["ref_rms"] > 0 and np.isfinite(e[k]["rms"])
        assert e[k]["rms"] < 1e-4 * e[k]["ref_rms"], (k, e[k])
    assert e["ssh"]["rms"] < e_raw["ssh"]["rms"], (e["ssh"], e_raw["ssh"])
    assert abs(d32[0] - d64[0]) <= 1e-5 * abs(d64[0])


def single_member(m, k, nsteps):
    with nemo2d.Model(int(m["jpi"][k]), int(m["jpj"][k]), dep_const=float(m["dep"][k]), rdt=RDT4,
                      cbfr=0.0015, visc=0.1) as s:
        s.set_tide(float(m["amp"][k]), float(m["period"][k]))
        s.record_checksums(nsteps)
        s.run(1, nsteps)
        r = s.retrieve()
        return {f: r[f] for f in ("sshn", "un", "vn")}, s.checksum_log(nsteps)


def test_ensemble_256_basins_3000_steps_fp64_and_fp32():
    mem = nemo2d.ensemble_members(NMEM, seed=7)
    side = mem["jpi"]
    order = np.argsort(side)
    sample = sorted({int(order[0]), int(order[-1]), int(order[NMEM // 2]), 17, 200})
    print("config-4 ensemble: %d members, %d cells, sides %d..%d" % (
        NMEM, int(np.sum(side.astype(np.int64) ** 2)), side.min(), side.max()), flush=True)
    states = {}
    for prec in (64, 32):
        with nemo2d.Ensemble(mem, rdt=RDT4, precision=prec) as e:
            choice, ms = e.tune(nsteps=10)
            print("precision %d: tuned tiling %d (%.3f ms/step)" % (prec, choice, ms), flush=True)
            e.run(1, STEPS4, graph=True)
            states[prec] = {k: e.retrieve(k) for k in range(0, NMEM, 8)}
            if prec == 64:
                for k in sample:
                    r = e.retrieve(k)
                    ref, log_ref = single_member(mem, k, STEPS4)
                    for f in ("sshn", "un", "vn"):
                        assert np.array_equal(r[f], ref[f]), (k, f)
                    log = e.checksum_log(k, STEPS4)
                    assert np.allclose(log, log_ref, rtol=1e-13, atol=0), k
    worst = {}
    for k, a in states[32].items():
        err = nemo2d.field_errors(a, states[64][k])
        for f in ("ssh", "u", "v"):
            assert np.isfinite(err[f]["rms"]) and err[f]["ref_rms"] > 0, (k, f)
            rel = err[f]["rms"] / err[f]["ref_rms"]
            worst[f] = max(worst.get(f, 0.0), rel)
    print("fp32 ensemble, worst member RMS error relative to the FP64 RMS:", worst)
    for f in ("ssh", "u", "v"):
        assert worst[f] < 1e-4, (f, worst[f])
