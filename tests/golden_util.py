"""Helpers shared by the oracle and GPU parity tests: grouping of the golden
reference runs and comparison rules."""
import json
import math
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def load_golden():
    with open(os.path.join(GOLDEN, "ref_checksums.json")) as f:
        return json.load(f)["runs"]


def trajectories(runs, max_cells=None):
    """Group golden runs that differ only in nitend: one trajectory each.

    Returns a list of (namelist-with-largest-nitend, {nitend: (ua, va)})."""
    groups = {}
    for r in runs:
        nml = dict(r["namelist"])
        if max_cells and (nml["jpiglo"] - 2) * (nml["jpjglo"] - 2) > max_cells:
            continue
        key = json.dumps({k: v for k, v in nml.items() if k != "nitend"}, sort_keys=True)
        g = groups.setdefault(key, [nml, {}])
        g[1][nml["nitend"]] = (r["ua"], r["va"])
        if nml["nitend"] > g[0]["nitend"]:
            g[0] = nml
    return [tuple(v) for _, v in sorted(groups.items())]


def e16_8(x):
    """Fortran E16.8 rendering (gocean_mod.F90:73-81), stripped."""
    x = float(x)
    if x == 0.0:
        return "0.00000000E+00"
    sign = "-" if x < 0 else ""
    mant, ex = ("%.7E" % abs(x)).split("E")
    e = int(ex) + 1
    digits = mant.replace(".", "")
    if abs(e) <= 99:
        return "%s0.%sE%+03d" % (sign, digits, e)
    return "%s0.%s%+04d" % (sign, digits, e)


def e16_8_close(x, ref):
    """True if x prints as ref in E16.8, or sits within a hair of the
    rounding boundary of ref's last digit (the reference sums in a different
    order; 2e-14 relative)."""
    if e16_8(x) == ref:
        return True
    r = float(ref.replace("-", "E-") if ("E" not in ref and ref.count("-") > 0) else ref)
    if r == 0.0:
        return x == 0.0
    ulp = 10.0 ** (math.floor(math.log10(abs(r))) - 7)
    return abs(abs(x - r) - 0.5 * ulp) <= 2e-14 * abs(r)


def rel_close(a, b, rtol=1e-12):
    """|a-b| <= rtol*max(|a|,|b|) elementwise; 0 == 0 passes."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a - b) <= rtol * np.maximum(np.abs(a), np.abs(b))


def load_traces(max_cells=None):
    """Per-step checksum traces of the reference (ref_step_checksums.npz):
    list of (namelist, array (nsteps, 4) of istp, sum|sshn|, sum|un|, sum|vn|)."""
    path = os.path.join(GOLDEN, "ref_step_checksums.npz")
    z = np.load(path)
    out = []
    for k, nml in enumerate(json.loads(str(z["namelists"]))):
        if max_cells and (nml["jpiglo"] - 2) * (nml["jpjglo"] - 2) > max_cells:
            continue
        out.append((nml, z["t%d" % k]))
    return out


def trace_id(t):
    nml = t[0]
    return "g%dx%d_n%d-%d_cb%s" % (nml["jpiglo"], nml["jpjglo"], nml["nit000"], nml["nitend"],
                                   nml["cbfr"])


def seq_sums(sshn, un, vn):
    """The reference's SUM(ABS(.)) of sshn, un, vn over (1:jpi,1:jpj) in its
    storage order (first index fastest), added one term at a time as its
    SUM does: arrays in this package's layout (T (jpj,jpi); U (jpj,jpi+1)
    from ji=0; V (jpj+1,jpi) from jj=0)."""
    return np.array([np.cumsum(np.abs(a.ravel()))[-1] if a.size else 0.0
                     for a in (sshn, un[:, 1:], vn[1:, :])])
