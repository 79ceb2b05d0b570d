"""GPU parity: the HIP library (through its C-ABI) against the C oracle and
the reference's golden checksums.

Bar: bitwise equality of every field for the same inputs (the library keeps
the serial model's operation order with FMA contraction off, and evaluates
sin/sqrt on the host with the same C library as the oracle); per-step
checksums within 1e-12 relative (both sides use compensated summation, so in
practice they differ by at most a few ulps)."""
import math
import os

import numpy as np
import pytest

import nemo2d
from golden_util import (e16_8, e16_8_close, load_golden, load_traces, rel_close, seq_sums,
                         trace_id, trajectories)
from oracle import Oracle

pytestmark = pytest.mark.gpu

FIELDS_T = ("sshn",)
FIELDS = ("sshn", "sshn_u", "sshn_v", "un", "vn")


def copy_grid_from_oracle(o):
    g = {k: np.array(o.arr(k)) for k in nemo2d.Model.GRID_KEYS}
    g["pt"] = np.array(o.pt)
    return g


def state_from_oracle(o):
    return {k: np.array(o.arr(k)) for k in FIELDS}


GRID_ARRAYS = ("e1t", "e2t", "e1u", "e2u", "e1v", "e2v", "e12t", "e12u", "e12v", "ht", "hu", "hv")


def assert_grid_on_device(m, g):
    """The device holds the grid as it was uploaded: every metric and depth
    array bit for bit, and the class words the host derives from the mask."""
    r = m.retrieve_grid()
    for k in GRID_ARRAYS:
        a, b = r[k], np.ascontiguousarray(g[k], dtype=np.float64)
        if not np.array_equal(a.view(np.uint64), b.view(np.uint64)):
            raise AssertionError("device %s differs from the uploaded array at %d points" % (
                k, int(np.count_nonzero(a.view(np.uint64) != b.view(np.uint64)))))
    cls = m.retrieve_classes()
    host = nemo2d.classify(g)
    if not np.array_equal(cls, host):
        raise AssertionError("device class words differ from the host's at %d points" %
                             int(np.count_nonzero(cls != host)))


def model_like(o, **kw):
    p = o.params
    g = copy_grid_from_oracle(o)
    m = nemo2d.Model(o.jpi, o.jpj, dx=p["dx"], dy=p["dy"], dep_const=p["dep_const"],
                     rdt=p["rdt"], cbfr=p["cbfr"], visc=p["visc"],
                     grid=g, state=state_from_oracle(o), **kw)
    assert_grid_on_device(m, g)
    return m


def assert_fields_equal(m, o, what=""):
    r = m.retrieve()
    for k in FIELDS:
        a, b = r[k], o.arr(k)
        if not np.array_equal(a.view(np.uint64), np.asarray(b).view(np.uint64)):
            bad = np.argwhere(a.view(np.uint64) != np.asarray(b).view(np.uint64))
            jj, ji = bad[0]
            raise AssertionError("%s %s differs at [%d,%d]: gpu %r oracle %r (%d points)" % (
                what, k, jj, ji, a[jj, ji], b[jj, ji], len(bad)))


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("jpi,jpj", [(1, 1), (2, 2), (1, 5), (5, 1), (3, 7), (16, 10), (65, 9),
                                     (128, 128)])
@pytest.mark.parametrize("path", ["fused", "stages", "stages_eager"])
def test_bitwise_vs_oracle_default_grid(jpi, jpj, path):
    """'stages' records the nine stage calls and runs them with next as one
    fused pass; 'stages_eager' runs a kernel per stage call."""
    nsteps = 120 if jpi * jpj < 5000 else 60
    o = Oracle(jpi, jpj)
    with model_like(o) as m:
        if path == "stages_eager":
            m.set_stage_fusion(False)
        m.record_checksums(nsteps)
        for istp in range(1, nsteps + 1):
            (m.step if path == "fused" else m.step_stages)(istp)
            o.step(istp)
            if istp in (1, 2, 3, 10, nsteps):
                assert_fields_equal(m, o, "%s step %d" % (path, istp))
        log = m.checksum_log(nsteps)
    o2 = Oracle(jpi, jpj)
    ref, _ = o2.run(1, nsteps)
    assert log.shape == ref.shape
    assert np.all(rel_close(log, ref, 1e-12)), np.max(np.abs(log - ref) / np.maximum(np.abs(ref), 1e-300))


def test_stage_by_stage_vs_oracle():
    """Every one of the nine calls, compared after the call, on a state that
    has been evolving (so every term is non-zero)."""
    o = Oracle(40, 24, cbfr=0.00015)
    o.run(1, 150)
    with model_like(o) as m:
        for istp in (151, 152):
            rt = istp * o.params["rdt"]
            m.continuity(); o.continuity()
            assert np.array_equal(m.retrieve()["ssha"], o.arr("ssha"))
            # Until the boundary stages run, the after-arrays still hold the
            # previous step's boundary values on the oracle side but the zero
            # of nemo2d_set_state on the GPU side (the first step of a model
            # starts from the serial model's zeroed ua/va): compare the
            # points the momentum stages write.
            m.momentum_u(); o.momentum_u()
            assert np.array_equal(m.retrieve()["ua"][:, 1:-1], o.arr("ua")[:, 1:-1])
            m.momentum_v(); o.momentum_v()
            assert np.array_equal(m.retrieve()["va"][1:-1, :], o.arr("va")[1:-1, :])
            m.bc_ssh(rt); o.bc_ssh(rt)
            assert np.array_equal(m.retrieve()["ssha"], o.arr("ssha"))
            m.bc_solid_u(); o.bc_solid_u()
            assert np.array_equal(m.retrieve()["ua"], o.arr("ua"))
            m.bc_solid_v(); o.bc_solid_v()
            assert np.array_equal(m.retrieve()["va"][1:, :], o.arr("va")[1:, :])
            m.bc_flather_u(); o.bc_flather_u()
            assert np.array_equal(m.retrieve()["ua"], o.arr("ua"))
            m.bc_flather_v(); o.bc_flather_v()
            assert np.array_equal(m.retrieve()["va"], o.arr("va"))
            m.next(); o.next()
            assert_fields_equal(m, o, "next %d" % istp)


def test_stage_calls_out_of_order_and_interrupted():
    """Stage calls that leave the serial model's order, or that another
    entry point interrupts (a retrieve, a tide change), run as called: the
    recorded calls are run first, stage by stage.  Each pattern gives the
    oracle's fields doing the same calls, bit for bit, with the recording
    on and off."""
    o = Oracle(48, 30, cbfr=0.00015)
    o.run(1, 60)
    start = state_from_oracle(o)
    rdt = o.params["rdt"]

    def step_calls(x, istp, pattern, extra):
        rt = istp * rdt
        calls = [x.continuity, x.momentum_u, x.momentum_v, lambda: x.bc_ssh(rt), x.bc_solid_u,
                 x.bc_solid_v, x.bc_flather_u, x.bc_flather_v]
        if pattern == "swap_solid":
            calls[4], calls[5] = calls[5], calls[4]
        if pattern == "momv_first":
            calls[1], calls[2] = calls[2], calls[1]
        for k, c in enumerate(calls):
            c()
            if extra and k == 3:
                extra()
        x.next()

    for pattern in ("swap_solid", "momv_first", "retrieve_mid", "in_order"):
        results = []
        for fusion in (True, False):
            oo = Oracle(48, 30, cbfr=0.00015)
            oo.run(1, 60)
            with model_like(oo) as m:
                m.set_stage_fusion(fusion)
                m.set_state(start)
                for istp in range(61, 71):
                    extra = (lambda: m.retrieve()) if pattern == "retrieve_mid" else None
                    step_calls(m, istp, pattern, extra)
                    step_calls(oo, istp, pattern, None)
                assert_fields_equal(m, oo, "%s fusion=%s" % (pattern, fusion))
                results.append(m.retrieve())
        for k in FIELDS:
            assert np.array_equal(results[0][k], results[1][k]), (pattern, k)


def test_stage_fusion_tide_change_mid_step():
    """A tide change between a recorded bc_ssh and next leaves the recorded
    call with the tide it was called with: the same fields as running each
    call at once."""
    o = Oracle(64, 40)
    o.run(1, 30)
    out = []
    for fusion in (True, False):
        with model_like(o) as m:
            m.set_stage_fusion(fusion)
            m.set_tide(1.0, 12.42)
            for istp in range(31, 41):
                rt = istp * o.params["rdt"]
                m.continuity(); m.momentum_u(); m.momentum_v(); m.bc_ssh(rt)
                m.set_tide(1.0 + 0.05 * (istp - 30), 12.42)
                m.bc_solid_u(); m.bc_solid_v(); m.bc_flather_u(); m.bc_flather_v(); m.next()
            out.append(m.retrieve())
    for k in FIELDS:
        assert np.array_equal(out[0][k], out[1][k]), k


def test_stage_timers_with_recorded_stage_calls():
    """Recorded stage calls count in their regions with no time; Next holds
    the fused pass.  Run at once, every region takes time."""
    o = Oracle(96, 64)
    for fusion in (True, False):
        with model_like(o) as m:
            m.set_stage_fusion(fusion)
            m.stage_timers(True)
            for istp in range(1, 6):
                m.step_stages(istp)
            t = m.stage_times()
            for r in ("Continuity", "Momentum", "BCs", "Next"):
                assert t[r][0] == 5, (fusion, r, t[r])
            assert t["Next"][1] > 0.0
            for r in ("Continuity", "Momentum", "BCs"):
                assert (t[r][1] == 0.0) if fusion else (t[r][1] > 0.0), (fusion, r, t[r])


def test_fused_and_stages_mix():
    """Fused and stage steps can be interleaved on one model."""
    o = Oracle(30, 20)
    with model_like(o) as m:
        for istp in range(1, 41):
            (m.step if istp % 3 else m.step_stages)(istp)
            o.step(istp)
        assert_fields_equal(m, o, "mixed")
        r = m.retrieve()
        assert np.array_equal(r["ua"], r["un"]) or np.array_equal(r["ua"], o.arr("ua"))


@pytest.mark.parametrize("variant", ["auto", "cell", "march16"])
def test_checksum_log_with_mixed_paths_and_restarts(variant):
    """The per-step log (fused steps leave block partials that the log read
    reduces; stage steps log at once) matches exact sums of the oracle's
    fields when fused and stage steps interleave and after set_state has
    discarded an unread recording; reading it twice gives the same log."""
    o = Oracle(150, 70)
    start = state_from_oracle(o)
    with model_like(o) as m:
        m.set_variant(variant)
        m.record_checksums(30)
        for istp in range(1, 11):
            m.step(istp)
        m.set_state(start)  # drops the 10 unread steps
        m.record_checksums(30)
        ref = []
        for istp in range(1, 31):
            (m.step_stages if istp % 4 == 0 else m.step)(istp)
            o.step(istp)
            # exact sums over (1:jpi, 1:jpj) of the oracle's fields
            ref.append([math.fsum(np.abs(o.arr("sshn")).ravel()),
                        math.fsum(np.abs(o.arr("un")[:, 1:]).ravel()),
                        math.fsum(np.abs(o.arr("vn")[1:, :]).ravel())])
        log1 = m.checksum_log(30)
        log2 = m.checksum_log(30)
        assert np.array_equal(log1, log2)
        assert np.allclose(log1, np.array(ref), rtol=1e-13, atol=0), np.abs(log1 / np.array(ref) - 1).max()


@pytest.mark.parametrize("traj", trajectories(load_golden(), max_cells=130 * 130),
                         ids=lambda t: "g%dx%d_n%d-%d_cb%s" % (t[0]["jpiglo"], t[0]["jpjglo"], t[0]["nit000"],
                                                             t[0]["nitend"], t[0]["cbfr"]))
def test_reference_golden_checksums(traj):
    """GPU per-step checksum log against the reference binary's ua/va lines."""
    nml, pts = traj
    n = nml["nitend"] - nml["nit000"] + 1
    with nemo2d.Model.from_namelist(nml) as m:
        m.record_checksums(n)
        m.run(nml["nit000"], nml["nitend"])
        log = m.checksum_log(n)
    for nitend, (ua, va) in pts.items():
        k = nitend - nml["nit000"]
        assert e16_8_close(log[k, 1], ua), (nitend, log[k, 1], ua)
        assert e16_8_close(log[k, 2], va), (nitend, log[k, 2], va)


def check_against_trace(nml, t, path="run", samples=None, rtol=1e-12):
    """The fields at the sample steps (default: the middle and last step),
    summed in the reference's order, equal to its printed 17 digits bit for
    bit, and the per-step GPU checksum log within rtol of those digits at
    every step, and within 1e-13 of the exact sums of its own fields at the
    sample steps."""
    n0, n1 = nml["nit000"], nml["nitend"]
    n = n1 - n0 + 1
    if samples is None:
        samples = [n0 + n // 2, n1]
    exact = {}
    with nemo2d.Model.from_namelist(nml) as m:
        m.record_checksums(n)
        a = n0
        for b in samples:
            if path == "run":
                m.run(a, b)
            else:
                for istp in range(a, b + 1):
                    m.step_stages(istp)
            a = b + 1
            r = m.retrieve()
            k = b - n0
            seq = seq_sums(r["sshn"], r["un"], r["vn"])
            if not np.array_equal(seq, t[k, 1:]):
                raise AssertionError("step %d: gpu %r reference %r" % (b, seq, t[k, 1:]))
            exact[k] = [float(np.sum(np.abs(x).astype(np.longdouble)))
                        for x in (r["sshn"], r["un"][:, 1:], r["vn"][1:, :])]
        if a <= n1:
            m.run(a, n1) if path == "run" else [m.step_stages(istp) for istp in range(a, n1 + 1)]
        log = m.checksum_log(n)
    assert log.shape == (n, 3)
    for k, e in exact.items():
        assert np.all(rel_close(log[k], e, 1e-13)), (n0 + k, log[k], e)
    rel = np.abs(log - t[:, 1:]) / np.maximum(np.maximum(np.abs(log), np.abs(t[:, 1:])), 1e-300)
    print("trace %s: largest relative difference of the log from the reference's digits %.3g at "
          "step %d" % (trace_id((nml, t)), rel.max(), n0 + int(np.argmax(rel.max(axis=1)))))
    ok = rel_close(log, t[:, 1:], rtol)
    if not ok.all():
        k, c = np.argwhere(~ok)[0]
        raise AssertionError("step %d column %d: gpu %r reference %r" % (n0 + k, c, log[k, c], t[k, 1 + c]))


@pytest.mark.parametrize("trace", load_traces(max_cells=130 * 130), ids=trace_id)
def test_reference_trace_every_step(trace):
    """The fused step against the reference's own per-step checksums."""
    check_against_trace(*trace)


def test_reference_trace_stage_path():
    nml, t = [x for x in load_traces(max_cells=130 * 130) if x[0]["jpiglo"] == 130][0]
    check_against_trace(nml, t, path="stages")


# ---- general masks and metrics ------------------------------------------------
def random_grid_oracle(jpi, jpj, seed, islands=True, open_edges="WEN", vary_metrics=True,
                       chained=False, vary_depths=False):
    rng = np.random.default_rng(seed)
    o = Oracle(jpi, jpj, dx=900.0, dy=1100.0, dep_const=80.0, rdt=15.0, cbfr=0.002, visc=5.0,
               default_grid=True)
    pt = o.pt
    if islands:
        land = rng.random((jpj, jpi)) < 0.08
        pt[1:jpj + 1, 1:jpi + 1][land] = 0
    if "W" in open_edges:
        pt[2:jpj // 2 + 1, 0] = -1
    if "E" in open_edges:
        pt[jpj // 3:jpj, jpi + 1] = -1
    if "N" in open_edges:
        pt[jpj + 1, 3:jpi - 1] = -1
    if chained:
        # water cell between two interior open cells: chained Flather points
        jj = jpj // 2
        pt[jj, 4:7] = [-1, 1, -1]
    if vary_metrics:
        for k in ("e1t", "e2t", "e1u", "e2u", "e1v", "e2v"):
            a = o.arr(k)
            a[:] = a * (1.0 + 0.2 * rng.random(a.shape))
        for k in ("ht", "hu", "hv"):
            a = o.arr(k)
            a[:] = 40.0 + 60.0 * rng.random(a.shape)
        for k in ("gphiu", "gphiv"):
            a = o.arr(k)
            a[:] = 30.0 + 20.0 * rng.random(a.shape)
        o.grid_areas()
    elif vary_depths == "derived":
        # hu, hv the means of adjacent T depths (as the seamount basins
        # define them): the default kernel forms them from ht on plain tiles
        ht, hu, hv = o.arr("ht"), o.arr("hu"), o.arr("hv")
        ht[:] = 40.0 + 60.0 * rng.random(ht.shape)
        hu[:, 1:jpi] = 0.5 * (ht[:, :-1] + ht[:, 1:])
        hu[:, 0], hu[:, jpi] = ht[:, 0], ht[:, -1]
        hv[1:jpj, :] = 0.5 * (ht[:-1, :] + ht[1:, :])
        hv[0, :], hv[jpj, :] = ht[0, :], ht[-1, :]
    elif vary_depths:
        for k in ("ht", "hu", "hv"):
            a = o.arr(k)
            a[:] = 40.0 + 60.0 * rng.random(a.shape)
    o.arr("sshn")[:] = 0.05 * rng.standard_normal(o.arr("sshn").shape)
    o.initialise()
    un, vn = o.arr("un"), o.arr("vn")
    un[:] = 0.02 * rng.standard_normal(un.shape)
    vn[:] = 0.02 * rng.standard_normal(vn.shape)
    return o


@pytest.mark.parametrize("seed,vary", [(1, True), (2, True), (3, True), (4, False), (5, False)])
@pytest.mark.parametrize("path", ["fused", "stages"])
def test_general_mask_and_metrics(seed, vary, path):
    """Islands and open edges on three sides; with per-point metrics, depths
    and latitudes (vary) or with the uniform ones (the marching kernel with
    Flather points on every edge)."""
    o = random_grid_oracle(137 if not vary else 37, 29, seed, vary_metrics=vary)
    with model_like(o) as m:
        for istp in range(1, 61):
            (m.step if path == "fused" else m.step_stages)(istp)
            o.step(istp)
        assert_fields_equal(m, o, "general seed %d" % seed)


def test_chained_flather_falls_back_exactly():
    o = random_grid_oracle(24, 20, 7, islands=False, vary_metrics=False, chained=True)
    with model_like(o) as m:
        for istp in range(1, 41):
            m.step(istp)
            o.step(istp)
        assert_fields_equal(m, o, "chained")


def test_mask_the_reference_cannot_run_is_rejected():
    g = nemo2d.default_grid(8, 6)
    g["pt"][3, 0] = 1      # water on the west ring: momentum would read un(-1, j)
    with pytest.raises(nemo2d.Nemo2dError) as e:
        nemo2d.Model(8, 6, grid=g)
    assert e.value.code == nemo2d.ERR_MASK
    g = nemo2d.default_grid(8, 6)
    g["pt"][2, 3] = 2
    with pytest.raises(nemo2d.Nemo2dError) as e:
        nemo2d.Model(8, 6, grid=g)
    assert e.value.code == nemo2d.ERR_MASK


def test_nonfinite_checksum_is_flagged():
    """A CFL-violating time step blows up; the checksum call reports it."""
    with nemo2d.Model(32, 32, rdt=400.0, visc=0.0) as m:
        m.run(1, 400)
        L = nemo2d.lib()
        out = np.zeros(3)
        rc = L.nemo2d_checksums(m._h, nemo2d._dp(out))
        assert rc == nemo2d.ERR_NONFINITE
        assert not np.all(np.isfinite(out))


def test_deterministic_repeat():
    a = nemo2d.Model(300, 200, cbfr=0.00015)
    b = nemo2d.Model(300, 200, cbfr=0.00015)
    a.record_checksums(200)
    b.record_checksums(200)
    a.run(1, 200)
    b.run(1, 200)
    la, lb = a.checksum_log(200), b.checksum_log(200)
    assert np.array_equal(la, lb)
    ra, rb = a.retrieve(), b.retrieve()
    for k in FIELDS:
        assert np.array_equal(ra[k], rb[k])


# ---- full size (4096^2) -------------------------------------------------------
def test_4096_first_steps_bitwise_vs_oracle():
    o = Oracle(4096, 4096)
    with nemo2d.Model(4096, 4096) as m:
        for istp in (1, 2, 3):
            m.step(istp)
            o.step(istp)
        assert_fields_equal(m, o, "4096 step 3")


def test_4096_reference_trace_every_step():
    """4096^2 x 1000 steps against the reference's own per-step sums: the
    fields at 10 steps from the first to the last, summed in the reference's
    order, bit for bit; the log within 1e-11 of the reference's digits at
    every step.  (The reference's SUM adds up to 16.7M terms one at a time
    and rounds at each: its digits stray from the exact sums of the same
    fields by up to 2.7e-12 relative (measured, step 992), while the log
    matches the exact sums to 1e-13; hence 1e-11 here and 1e-12 on the
    smaller grids, where the largest difference measured is 8.7e-14.)"""
    big = [x for x in load_traces() if x[0]["jpiglo"] == 4098]
    assert big, "no 4096^2 trace in tests/golden"
    check_against_trace(*big[0], samples=[1, 2, 7, 10, 50, 100, 250, 500, 750, 1000], rtol=1e-11)


def test_4096_reference_checksums_and_stage_path():
    runs = [r for r in load_golden() if r["namelist"]["jpiglo"] == 4098]
    if not runs:
        pytest.skip("no 4096^2 reference checksums in tests/golden")
    nml = max((r["namelist"] for r in runs), key=lambda n: n["nitend"])
    n = nml["nitend"]
    with nemo2d.Model.from_namelist(nml) as m:
        m.record_checksums(n)
        m.run(1, n)
        log = m.checksum_log(n)
    for r in runs:
        k = r["namelist"]["nitend"] - 1
        assert e16_8_close(log[k, 1], r["ua"]), (k + 1, log[k, 1], r["ua"])
        assert e16_8_close(log[k, 2], r["va"]), (k + 1, log[k, 2], r["va"])
    # the stage path gives the same fields bit for bit over 50 steps (and
    # checksums to rounding: the two paths reduce in different orders)
    with nemo2d.Model.from_namelist(nml) as s, nemo2d.Model.from_namelist(nml) as f:
        s.record_checksums(50)
        for istp in range(1, 51):
            s.step_stages(istp)
        f.run(1, 50)
        assert np.allclose(s.checksum_log(50), log[:50], rtol=1e-13, atol=0)
        rs, rf = s.retrieve(), f.retrieve()
        for k in FIELDS:
            assert np.array_equal(rs[k], rf[k]), k


# ---- the Fortran driver -------------------------------------------------------
NAMELIST = """&namctl
   jpiglo = {g}
   jpjglo = {g}
   jphgr_msh = 1
   dx = 1000.0
   dy = 1000.0
   dep_const = 100.0
   nit000 = 1
   nitend = {n}
   irecord = {irec}
   rdt = 20.0
   cbfr = 0.0015
   visc = 0.1
/
&namhip
   fused = {fused}
   stage_fusion = {sf}
   write_output = .true.
   async_output = {asy}
   csfile = 'cs.txt'
/
"""


@pytest.mark.parametrize("fused,asy,sf", [(".true.", ".false.", ".true."),
                                         (".false.", ".false.", ".true."),
                                         (".false.", ".false.", ".false."),
                                         (".true.", ".true.", ".true."),
                                         (".false.", ".true.", ".true."),
                                         (".false.", ".true.", ".false.")])
def test_fortran_driver(tmp_path, fused, asy, sf):
    """The driver's three step paths: the fused step, the nine calls run as
    one fused step (stage_fusion, the default) and the nine calls one kernel
    each; with snapshots written in line or behind the steps."""
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "nemolite2d-cuda_amd", "bin", "nemolite2d_hip")
    g, n = 66, 300
    (tmp_path / "namelist").write_text(NAMELIST.format(g=g, n=n, irec=150, fused=fused, asy=asy,
                                                       sf=sf))
    p = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    o = Oracle(g - 2, g - 2)
    ref, _ = o.run(1, n)
    lines = [l for l in p.stdout.splitlines() if "checksum =" in l]
    assert lines[0].split("=")[1].strip() == e16_8(ref[-1, 1])
    assert lines[1].split("=")[1].strip() == e16_8(ref[-1, 2])
    # timing rows as the reference's dl_timer report lays them out (what its
    # benchmark extractor splits): label, count, total, average, avg/repeat, err
    rows = {l.split()[0]: l.split() for l in p.stdout.splitlines()
            if l.startswith(("Time-stepping", "Continuity", "Momentum", "BCs", "Next", "Fused"))}
    assert len(rows["Time-stepping"]) == 6 and float(rows["Time-stepping"][2]) > 0
    names = ("Fused",) if fused == ".true." else ("Continuity", "Momentum", "BCs", "Next")
    for r in names:
        assert int(rows[r][-5 if r == "Fused" else 1]) == n, rows[r]
    cs = np.loadtxt(tmp_path / "cs.txt")
    assert cs.shape == (n, 4)
    assert np.all(rel_close(cs[:, 1:], ref, 1e-12))
    # snapshots at istp = 0, 150, 300 in the serial model's format
    for k in (0, 150, 300):
        d = np.loadtxt(tmp_path / ("go2d_%05d.dat" % k))
        assert d.shape == ((g - 2) ** 2, 6)
    d = np.loadtxt(tmp_path / "go2d_00300.dat")
    assert np.allclose(d[:, 3], o.arr("sshn").ravel(), rtol=1e-6, atol=1e-12)
    # the text itself: the oracle's state through the library's formatter
    # (pinned byte for byte to a reference file on the CPU)
    txt = nemo2d.format_go2d(o.arr("xt"), o.arr("yt"), o.arr("ht"), o.arr("sshn"), o.arr("un"),
                             o.arr("vn"))
    assert (tmp_path / "go2d_00300.dat").read_bytes() == txt


GRID_NAMELIST = """&namctl
   jpiglo = {gi}
   jpjglo = {gj}
   jphgr_msh = {msh}
   dx = {dx}
   dy = {dy}
   dep_const = {dep}
   nit000 = 1
   nitend = {n}
   irecord = 100000
   rdt = {rdt}
   cbfr = {cbfr}
   visc = {visc}
/
&namhip
   fused = {fused}
   write_output = .false.
   csfile = 'cs.txt'
   grid_file = 'grid.txt'
/
"""


def run_driver(d, **kw):
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "nemolite2d-cuda_amd", "bin", "nemolite2d_hip")
    d.mkdir(exist_ok=True)
    (d / "namelist").write_text(GRID_NAMELIST.format(**kw))
    p = subprocess.run([exe], cwd=d, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    return p.stdout, np.loadtxt(d / "cs.txt")


def test_fortran_driver_grid_from_file(tmp_path):
    """jphgr_msh = 0: the driver reads the grid from a file.  The reference
    basin written to a file gives the built-in grid's checksums bit for bit;
    a grid with islands, open edges on three sides and varying metrics,
    depths and latitudes gives the oracle's checksums on that grid, through
    the fused step and through the nine stage calls."""
    base = dict(gi=42, gj=34, dx=1000.0, dy=1000.0, dep=100.0, n=120, rdt=20.0, cbfr=0.0015, visc=0.1,
                fused=".true.")
    (tmp_path / "a").mkdir()
    nemo2d.write_grid_file(str(tmp_path / "a" / "grid.txt"), nemo2d.default_grid(40, 32))
    out_a, cs_a = run_driver(tmp_path / "a", msh=0, **base)
    out_b, cs_b = run_driver(tmp_path / "b", msh=1, **base)
    assert np.array_equal(cs_a, cs_b)
    assert [l for l in out_a.splitlines() if "checksum =" in l] == \
        [l for l in out_b.splitlines() if "checksum =" in l]
    o = random_grid_oracle(37, 29, 3, vary_metrics=True)
    p = o.params
    (tmp_path / "c").mkdir()
    nemo2d.write_grid_file(str(tmp_path / "c" / "grid.txt"), copy_grid_from_oracle(o))
    # the driver starts from rest on this grid
    for k in ("sshn", "sshn_u", "sshn_v", "un", "vn"):
        o.arr(k)[:] = 0.0
    n = 80
    ref, _ = o.run(1, n)
    for fused in (".true.", ".false."):
        out, cs = run_driver(tmp_path / "c", msh=0, gi=39, gj=31, dx=p["dx"], dy=p["dy"],
                             dep=p["dep_const"], n=n, rdt=p["rdt"], cbfr=p["cbfr"], visc=p["visc"],
                             fused=fused)
        assert cs.shape == (n, 4)
        assert np.all(rel_close(cs[:, 1:], ref, 1e-12)), fused
        lines = [l for l in out.splitlines() if "checksum =" in l]
        assert lines[0].split("=")[1].strip() == e16_8(ref[-1, 1])
        assert lines[1].split("=")[1].strip() == e16_8(ref[-1, 2])


def test_snapshots_behind_steps_equal_synchronous_state(tmp_path):
    """snapshot_begin queues the copy and the file; the steps queued after
    it do not disturb it (the buffer it copies is next written two steps
    later, which waits for the copy)."""
    o = Oracle(150, 90)
    with model_like(o) as m:
        m.set_output_grid(o.arr("xt"), o.arr("yt"), o.arr("ht"))
        for istp in range(1, 13):
            m.step(istp)
            o.step(istp)
            if istp in (3, 4, 9):
                m.snapshot_begin(istp, str(tmp_path))
                want = nemo2d.format_go2d(o.arr("xt"), o.arr("yt"), o.arr("ht"), o.arr("sshn"),
                                          o.arr("un"), o.arr("vn"))
                (tmp_path / ("want_%05d.dat" % istp)).write_bytes(want)
        m.snapshot_wait()
        for istp in (3, 4, 9):
            assert (tmp_path / ("go2d_%05d.dat" % istp)).read_bytes() == \
                (tmp_path / ("want_%05d.dat" % istp)).read_bytes(), istp
        assert_fields_equal(m, o, "after snapshots")


@pytest.mark.parametrize("variant", ["auto", "march12_grd", "march8_grd", "march12_shd", "cell"])
@pytest.mark.parametrize("islands", [False, True])
@pytest.mark.parametrize("depths", [True, "derived"])
def test_per_point_depths_bitwise_vs_oracle(variant, islands, depths):
    """Uniform metrics, varying depths, the reference basin's open south edge
    (and inner walls): the marching kernel with per-point depths (the
    default for such large grids, named here since this one is small)
    against the oracle; 'derived': hu and hv the means of adjacent T depths,
    which that kernel forms from ht."""
    o = random_grid_oracle(203, 61, 21 + islands, islands=islands, open_edges="", vary_metrics=False,
                           vary_depths=depths)
    with model_like(o) as m:
        m.set_variant(variant)
        m.record_checksums(25)
        for istp in range(1, 26):
            m.step(istp)
            o.step(istp)
            if istp % 5 == 0:
                assert_fields_equal(m, o, "%s step %d" % (variant, istp))
        assert np.all(np.isfinite(m.checksum_log(25)))


@pytest.mark.parametrize("variant", ["auto", "march12_rc", "march8_rc"])
@pytest.mark.parametrize("islands", [False, True])
def test_sshn_only_kernel_bitwise_vs_oracle(variant, islands):
    """The fused kernel that reads sshn, un and vn only (sshn_u and sshn_v
    recomputed from sshn with next's expressions, the stored value kept where
    next leaves a point alone, original/nemolite2d.f90:735-762): uniform
    statics, the reference basin's open edge, inner walls from islands;
    interleaved with stage steps and with state reads (both fill the
    interpolated arrays in first)."""
    o = random_grid_oracle(203, 61, 31 + islands, islands=islands, open_edges="", vary_metrics=False)
    with model_like(o) as m:
        m.set_variant(variant)
        m.record_checksums(40)
        for istp in range(1, 41):
            (m.step_stages if istp % 13 == 0 else m.step)(istp)
            o.step(istp)
            if istp % 7 == 0:
                assert_fields_equal(m, o, "%s step %d" % (variant, istp))
        assert_fields_equal(m, o, "%s end" % variant)
        assert np.all(np.isfinite(m.checksum_log(40)))


@pytest.mark.parametrize("variant", ["auto", "march12_rc", "march8_rc"])
def test_sshn_only_kernel_after_a_state_with_other_interpolations(variant):
    """A state whose sshn_u/sshn_v are not sshn's interpolations (set by the
    caller) must be stepped from the stored values: the kernel that recomputes
    them is used only once a step has made them so."""
    o = Oracle(150, 70)
    rng = np.random.default_rng(5)
    for k in ("sshn", "sshn_u", "sshn_v"):
        a = o.arr(k)
        a[:] = 0.01 * rng.standard_normal(a.shape)
    with model_like(o) as m:
        m.set_variant(variant)
        for istp in range(1, 6):
            m.step(istp)
            o.step(istp)
            assert_fields_equal(m, o, "%s step %d" % (variant, istp))


@pytest.mark.parametrize("variant", ["auto", "march12_rc", "march12_dpf", "depths"])
def test_tiny_values_take_the_ieee_divisions(variant):
    """Numerators below 2^-960 (as in the subnormal band that trails the wave
    front of a run from rest) make the kernels that divide by reciprocals
    fall back to IEEE division (per division, or by redoing the row): a state
    full of such values, and of zeros, must still match the oracle bit for
    bit.  'depths': the same on a grid with per-point depths (the marching
    kernel for such grids)."""
    if variant == "depths":
        o = random_grid_oracle(300, 70, 44, islands=False, open_edges="", vary_metrics=False,
                               vary_depths="derived")
        variant = "march12_grd"
    else:
        o = Oracle(300, 70)
    rng = np.random.default_rng(9)
    for k, scale in (("sshn", 1e-300), ("un", 1e-305), ("vn", 1e-310)):
        a = o.arr(k)
        a[:] = rng.standard_normal(a.shape) * scale * (rng.random(a.shape) < 0.6)
    with model_like(o) as m:
        m.set_variant(variant)
        for istp in range(1, 9):
            m.step(istp)
            o.step(istp)
            assert_fields_equal(m, o, "%s step %d" % (variant, istp))


def test_default_kernel_by_grid_size():
    """The kernel the default picks, as the library reports it: one thread
    per cell (Flather inline) on the 128^2 basin, then the sshn-only kernel
    (from the first step: the state from rest passes set_state's
    interpolation check) with 8-row segments up to 3M cells and 12-row ones
    beyond; on a seamount grid the per-point-depth marching kernel, with
    8-row segments up to 10M cells."""
    cases = [(128, 128, None, ["cell_flin", "cell_flin"]),
             (1024, 1024, None, ["march8_rc", "march8_rc"]),
             (2000, 2000, None, ["march12_rc", "march12_rc"]),
             (1200, 1200, "seamount", ["march8_grd", "march8_grd"]),
             (3200, 3200, "seamount", ["march12_grd", "march12_grd"])]
    for ni, nj, kind, want in cases:
        grid = nemo2d.seamount_grid(ni, nj, seed=2)[0] if kind else None
        with nemo2d.Model(ni, nj, grid=grid) as m:
            assert m.last_variant() is None
            got = []
            for istp in (1, 2):
                m.step(istp)
                got.append(m.last_variant())
            assert got == want, (ni, nj, kind, got)


def test_first_step_kernel_follows_the_interpolation_check():
    """nemo2d_set_state checks whether the given sshn_u/sshn_v are next's
    interpolations of sshn bit for bit.  An evolved state (written by next)
    passes: the first step already takes the sshn-only kernel.  The same
    state with one sshn_u value one ulp off fails: the first step takes the
    five-array kernel, which reads the stored value.  Both runs bitwise equal
    to the oracle from the same state."""
    for perturb in (False, True):
        o = Oracle(1500, 1500)
        o.run(1, 3)
        if perturb:
            su = o.arr("sshn_u")
            su[700, 700] = np.nextafter(su[700, 700], np.inf)
        with model_like(o) as m:
            kernels = []
            for istp in (4, 5):
                m.step(istp)
                o.step(istp)
                kernels.append(m.last_variant())
            assert kernels == (["march12_dpfb", "march8_rc"] if perturb else ["march8_rc"] * 2), kernels
            assert_fields_equal(m, o, "perturbed" if perturb else "evolved")


@pytest.mark.parametrize("variant", [v for v in nemo2d.VARIANTS if v != "auto"])
def test_every_kernel_variant_is_bitwise_equal(variant):
    """Default grid (uniform statics: all variants apply; every divisor d
    has |d*RN(1/d) - 1| <= 2^-54, so the reciprocal divisions take one
    correction), the same basin 150 m deep (1/150 fails that test: two
    corrections in the uniform kernel; the per-point-depth kernel divides
    by the depths with IEEE division anyway), one with dx = dy = 1700 m
    (1/1700 fails it: the per-point-depth kernel's metric divisions stay
    IEEE) and a general mask with uniform metrics (Flather points on three
    edges)."""
    for o in (Oracle(300, 70), Oracle(300, 70, dep_const=150.0), Oracle(300, 70, dx=1700.0, dy=1700.0),
              random_grid_oracle(137, 45, 11, vary_metrics=False)):
        with model_like(o) as m:
            m.set_variant(variant)
            m.record_checksums(30)
            for istp in range(1, 31):
                m.step(istp)
                o.step(istp)
            assert_fields_equal(m, o, variant)
            log = m.checksum_log(30)
            assert np.all(np.isfinite(log))
