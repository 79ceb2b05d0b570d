"""Pins the C oracle (oracle/nemo2d_oracle.c) before it is trusted as the
checker of the GPU path.

* golden vectors from the reference binary (tests/golden/, made by
  tests/golden/make_golden.py from oracle/_ref): end-of-run ua/va checksums
  for many nitend values, grids and parameters, and decimated go2d_*.dat
  snapshots;
* closed forms and invariants that a dropped term, a wrong sign or a wrong
  index would break.
"""
import json
import math
import os

import numpy as np
import pytest

from golden_util import (GOLDEN, e16_8, e16_8_close, load_golden, load_traces, rel_close,
                         seq_sums, trace_id, trajectories)
from oracle import Oracle, fortran_e16_8, tide

RUNS = load_golden()


def test_golden_present():
    assert len(RUNS) >= 80


@pytest.mark.parametrize("traj", trajectories(RUNS, max_cells=130 * 130),
                         ids=lambda t: "g%dx%d_n%d-%d_cb%s_vi%s" % (
                             t[0]["jpiglo"], t[0]["jpjglo"], t[0]["nit000"], t[0]["nitend"],
                             t[0]["cbfr"], t[0]["visc"]) if isinstance(t, tuple) else str(t))
def test_oracle_matches_reference_checksums(traj):
    nml, pts = traj
    o = Oracle.from_namelist(nml)
    sums, _ = o.run(nml["nit000"], nml["nitend"])
    for nitend, (ua, va) in sorted(pts.items()):
        k = nitend - nml["nit000"]
        assert e16_8_close(sums[k, 1], ua), (nitend, e16_8(sums[k, 1]), ua)
        assert e16_8_close(sums[k, 2], va), (nitend, e16_8(sums[k, 2]), va)


def test_fortran_e16_8_format():
    assert fortran_e16_8(10.917447449) == "  0.10917447E+02"
    assert fortran_e16_8(0.0) == "  0.00000000E+00"
    assert fortran_e16_8(1e-300) == "  0.10000000-299"
    assert e16_8(140.67051) == "0.14067051E+03"


def _print_tol(ref):
    """Half a unit of the last digit printed by '(6e20.7e4)'."""
    a = np.abs(ref)
    e = np.floor(np.log10(np.where(a > 0, a, 1.0))) + 1
    return 0.5e-7 * 10.0 ** e * (1 + 1e-9)


@pytest.mark.parametrize("tag", ["128", "40x30"])
def test_oracle_matches_reference_snapshots(tag):
    z = np.load(os.path.join(GOLDEN, "ref_fields_%s.npz" % tag))
    nml = json.loads(str(z["namelist"]))
    stride = int(z["stride"])
    o = Oracle.from_namelist(nml)
    prev = 0
    for k, s in enumerate(z["steps"]):
        if s > prev:
            o.run(prev + 1, int(s))
        prev = int(s)
        sshn, un, vn = o.arr("sshn"), o.arr("un"), o.arr("vn")
        mine = np.stack([sshn, 0.5 * (un[:, :-1] + un[:, 1:]), 0.5 * (vn[:-1, :] + vn[1:, :])], -1)
        mine = mine[::stride, ::stride]
        ref = z["fields"][k]
        assert mine.shape == ref.shape
        err = np.abs(mine - ref)
        assert np.all(err <= _print_tol(ref)), (tag, s, float(err.max()))


# ---- closed forms ------------------------------------------------------------

def test_rest_state_first_step():
    """Step 1 from rest: continuity gives 0, the south row is clamped to the
    tide at rtime=rdt, velocities stay 0 (the pressure gradient sees the old
    ssh, which is 0)."""
    o = Oracle(20, 12, rdt=20.0)
    o.step(1)
    ssh = o.arr("sshn")
    t = tide(20.0)
    assert t == 0.2 * math.sin(2.0 * 3.14159 / (12.42 * 3600.0) * 20.0)
    assert np.all(ssh[0, :] == t)
    assert np.all(ssh[1:, :] == 0.0)
    assert np.all(o.arr("un") == 0.0) and np.all(o.arr("vn") == 0.0)
    c = o.checksums()
    assert c[0] == pytest.approx(20 * abs(t), rel=1e-15)
    # next: interior sshn_v rows interpolate between row 1 (tide) and row 2 (0)
    sv = o.arr("sshn_v")
    assert np.all(sv[1, :] == 0.5 * (1e6 * t + 1e6 * 0.0) / 1e6)
    assert np.all(sv[0, :] == 0.0)           # row 0 is never updated
    su = o.arr("sshn_u")
    assert np.all(su[0, :] == t)             # uniform row: average = copy = t


def test_second_step_hand_derived():
    """Step 2: only the first row has ssh; the pressure gradient pushes v at
    V row 1 north: va(ji,1) = rdt*hpg/e12v/hv/(1+cbfr*rdt), hpg from the
    ssh jump t1 between rows 1 and 2 (all other terms vanish at rest)."""
    rdt, cbfr, g, h = 20.0, 0.0015, 9.80665, 100.0
    o = Oracle(16, 10, rdt=rdt, cbfr=cbfr, visc=0.1)
    o.step(1)
    t1 = tide(rdt)
    sv1 = 0.5 * (1e6 * t1 + 1e6 * 0.0) / 1e6
    o.step(2)
    va = o.arr("vn")
    hpg = -g * (h + sv1) * 1000.0 * (0.0 - t1)
    expect = (0.0 * (h + sv1) + rdt * (0.0 + 0.0 + 0.0 + hpg) / 1e6) / (h + 0.0) / (1.0 + cbfr * rdt)
    assert np.all(va[1, :] == expect)
    # Flather row 0 = row 1 + sqrt(g/h) * (sshn_v(0) - sshn_v(1)) with old sshn_v
    assert np.allclose(va[0, :], expect + math.sqrt(g / h) * (0.0 - sv1), rtol=0, atol=0)
    assert np.all(va[-1, :] == 0.0)          # bc_solid_v zeroes row jpj
    un = o.arr("un")
    assert np.all(un[:, 0] == 0.0) and np.all(un[:, -1] == 0.0)   # bc_solid_u


def test_continuity_mass_budget():
    """sum(e12t*(ssha-sshn)) = rdt * net inflow through the open south edge,
    for every step of a nonlinear run (continuity is a flux divergence; the
    closed walls carry no flux; clamped cells are excluded)."""
    o = Oracle(24, 18, rdt=20.0)
    o.run(1, 40)
    sshn = o.arr("sshn").copy()
    o.continuity()
    ssha = o.arr("ssha")
    e12t = o.arr("e12t")
    hv, sv, vn = o.arr("hv"), o.arr("sshn_v"), o.arr("vn")
    hu, su, un = o.arr("hu"), o.arr("sshn_u"), o.arr("un")
    lhs = np.sum(e12t * (ssha - sshn))
    inflow = np.sum((sv[0, :] + hv[0, :]) * vn[0, :]) - np.sum((sv[-1, :] + hv[-1, :]) * vn[-1, :])
    inflow += np.sum((su[:, 0] + hu[:, 0]) * un[:, 0]) - np.sum((su[:, -1] + hu[:, -1]) * un[:, -1])
    assert lhs == pytest.approx(20.0 * inflow, rel=1e-9, abs=1e-9)


def test_uniform_ssh_interpolation():
    """next: a uniform new ssh gives sshn_u = sshn_v = that value everywhere
    they are updated (average and both copies)."""
    o = Oracle(9, 7)
    o.arr("ssha")[:] = 0.125
    o.arr("ua")[:] = 0.0
    o.arr("va")[:] = 0.0
    o.next()
    assert np.all(o.arr("sshn") == 0.125)
    assert np.all(o.arr("sshn_u") == 0.125)
    assert np.all(o.arr("sshn_v")[1:, :] == 0.125)


def test_uniform_pressure_gradient():
    """A uniform ssh slope in x with everything else at rest gives the
    analytic pressure-gradient velocity at interior U points."""
    rdt, cbfr, g, h, dx = 20.0, 0.0015, 9.80665, 100.0, 1000.0
    o = Oracle(12, 8, rdt=rdt, cbfr=cbfr, visc=0.0)
    slope = 1e-4
    x = np.arange(1, 13, dtype=np.float64)
    o.arr("sshn")[:] = slope * x[None, :]
    o.arr("sshn_u")[:] = 0.0
    o.momentum_u()
    ua = o.arr("ua")[:, 1:12]
    expect = rdt * (-g * h * 1000.0 * (slope * 1.0 + 0 * x[1:][None, :]) + 0.0) / 1e6 / h / (1 + cbfr * rdt)
    # advection/viscosity vanish with un = 0; Coriolis vanishes with vn = 0
    assert np.allclose(ua, expect, rtol=1e-12, atol=0)


def test_reflection_symmetry_without_coriolis():
    """With the Coriolis term off (gphi = 0) the basin is symmetric about its
    vertical mid-line: u is odd and ssh, v are even under x -> -x."""
    o = Oracle(10, 8)
    o.arr("gphiu")[:] = 0.0
    o.arr("gphiv")[:] = 0.0
    o.run(1, 60)
    ssh, un, vn = o.arr("sshn"), o.arr("un"), o.arr("vn")
    assert np.allclose(ssh, ssh[:, ::-1], rtol=1e-12, atol=1e-18)
    assert np.allclose(vn, vn[:, ::-1], rtol=1e-12, atol=1e-18)
    assert np.allclose(un, -un[:, ::-1], rtol=1e-12, atol=1e-18)
    # and with Coriolis on, the symmetry is broken (the term is not dropped)
    o2 = Oracle(10, 8)
    o2.run(1, 60)
    assert not np.allclose(o2.arr("sshn"), o2.arr("sshn")[:, ::-1], rtol=1e-9, atol=0)


# ---- per-step checksums of the reference at full precision --------------------
TRACES = load_traces(max_cells=130 * 130)


def test_traces_present():
    assert len(TRACES) >= 8
    assert {t[0]["cbfr"] for t in TRACES if t[0]["jpiglo"] == 130} == {0.0015, 0.00015}


@pytest.mark.parametrize("trace", TRACES, ids=trace_id)
def test_oracle_per_step_checksums_equal_reference_trace(trace):
    """Every step: the oracle's fields, summed as the reference's SUM adds
    them (in storage order), give the reference's printed 17 digits bit for
    bit; its compensated checksums are within 1e-12 relative."""
    nml, t = trace
    o = Oracle.from_namelist(nml)
    n0, n1 = nml["nit000"], nml["nitend"]
    assert np.array_equal(t[:, 0], np.arange(n0, n1 + 1))
    for k, istp in enumerate(range(n0, n1 + 1)):
        o.step(istp)
        seq = seq_sums(o.arr("sshn"), o.arr("un"), o.arr("vn"))
        if not np.array_equal(seq, t[k, 1:]):
            raise AssertionError("step %d: oracle %r reference %r" % (istp, seq, t[k, 1:]))
        assert np.all(rel_close(o.checksums(), t[k, 1:])), istp


def test_sequential_sums_sensitivity():
    """What the bitwise trace comparison can see: any one value off by one
    ulp of the sum, or every value off by 1e-14 relative.  (Every value one
    ulp off is below a 16384-term sum's own last bit: the fields themselves
    are pinned bit for bit to the oracle by the GPU parity tests; the trace
    pins the oracle's per-step sums to the reference's 17 digits.)"""
    nml, t = [x for x in TRACES if x[0]["jpiglo"] == 130][0]
    o = Oracle.from_namelist(nml)
    for istp in range(1, 301):
        o.step(istp)
    s, u, v = np.array(o.arr("sshn")), np.array(o.arr("un")), np.array(o.arr("vn"))
    assert np.array_equal(seq_sums(s, u, v), t[299, 1:])
    assert seq_sums(s * (1 + 1e-14), u, v)[0] != t[299, 1]
    assert seq_sums(s, u, v * (1 + 1e-14))[2] != t[299, 3]
    s2 = s.copy()
    s2[40, 50] += np.spacing(t[299, 1])
    assert seq_sums(s2, u, v)[0] != t[299, 1]


def test_python_initial_state_equals_oracle_initialise():
    """nemo2d.initial_state (the host side of the driver's initialisation,
    original/nemolite2d.f90:324-371) against the oracle's, on varying areas
    and a random ssh: bit for bit."""
    import nemo2d
    o = Oracle(23, 17)
    rng = np.random.default_rng(1)
    o.arr("e12t")[:] = 1e6 * (1 + 0.3 * rng.random(o.arr("e12t").shape))
    o.arr("sshn")[:] = rng.standard_normal((17, 23))
    o.initialise()
    g = {k: np.array(o.arr(k)) for k in ("e12t", "e12u", "e12v", "e1t")}
    st = nemo2d.initial_state(g, np.array(o.arr("sshn")))
    for k in ("sshn_u", "sshn_v"):
        assert np.array_equal(st[k], o.arr(k)), k


def test_oracle_4096_first_steps_equal_reference_trace():
    """The 4096^2 trace's first steps (the full 1000 are checked on the GPU
    path, which equals the oracle bit for bit)."""
    big = [t for t in load_traces() if t[0]["jpiglo"] == 4098]
    assert big, "no 4096^2 trace in tests/golden"
    nml, t = big[0]
    assert t.shape == (1000, 4)
    o = Oracle.from_namelist(nml)
    for istp in (1, 2, 3):
        o.step(istp)
        seq = seq_sums(o.arr("sshn"), o.arr("un"), o.arr("vn"))
        assert np.array_equal(seq, t[istp - 1, 1:]), (istp, seq, t[istp - 1, 1:])
