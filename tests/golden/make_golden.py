#!/usr/bin/env python3
"""Generate golden vectors from the reference NEMOLite2D binary.

The reference's serial model (original/nemolite2d.f90) is built from its own
source files by ``oracle/ref.mk`` into ``oracle/_ref/nemolite2d_ref``.  This
script runs that binary (never anything from the reference tree itself) on
namelists it writes, in scratch directories under ``oracle/_ref/runs/``, and
records only numbers:

* ``ref_checksums.json`` -- the two lines the reference prints at the end of
  a run, ``ua checksum = E16.8`` and ``va checksum = E16.8``
  (original/nemolite2d.f90:82-85), for runs of ``nitend`` steps.  Running the
  model to several ``nitend`` values turns the end-of-run checksum into a
  per-step trajectory sample.
* ``ref_fields_<case>.npz`` -- a decimated sample of the ``go2d_NNNNN.dat``
  snapshots (x, y, depth, ssh, u at T, v at T written ``(6e20.7e4)`` by
  original/nemolite2d.f90:773-799) for point-wise checks.
* ``ref_go2d_10x10_00012.dat`` -- one such snapshot file byte for byte (a
  10x10 basin after 12 steps), the expected output of the library's
  snapshot writer on the same fields.
* ``ref_step_checksums.npz`` -- per-step checksums at full precision from
  ``oracle/_ref/nemolite2d_ref_trace``, the same model with one added
  statement that prints SUM(ABS(.)) of sshn, un and vn over (1:jpi,1:jpj)
  after every step (ES24.16, 17 significant digits, which a double round-
  trips exactly).  Fortran's SUM adds the section in storage order, so a
  sequential sum of the same fields in that order equals these values bit
  for bit.

Usage (from the repository root)::

    python tests/golden/make_golden.py            # 128^2 and tiny grids
    python tests/golden/make_golden.py --big      # adds 4096^2 (about 30 min)
    python tests/golden/make_golden.py --trace    # per-step checksums (add --big: 4096^2)
"""
import argparse
import json
import os
import re
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REF_EXE = os.path.join(ROOT, "oracle", "_ref", "nemolite2d_ref")
RUNS = os.path.join(ROOT, "oracle", "_ref", "runs")
HERE = os.path.dirname(os.path.abspath(__file__))

NAMELIST = """&namctl
   jpiglo = {jpiglo}
   jpjglo = {jpjglo}
   jphgr_msh = 1
   dx = {dx}
   dy = {dy}
   dep_const = {dep}
   nit000 = {nit000}
   nitend = {nitend}
   irecord = {irecord}
   rdt = {rdt}
   cbfr = {cbfr}
   visc = {visc}
/
"""

LINE = re.compile(r"^\s*(ua|va) checksum =\s*(\S+)\s*$")
TRACE_EXE = os.path.join(ROOT, "oracle", "_ref", "nemolite2d_ref_trace")
TRACE_LINE = re.compile(r"^\s*step checksums\s+(-?\d+)\s+(\S+)\s+(\S+)\s+(\S+)\s*$")


def case_name(c):
    return ("g{jpiglo}x{jpjglo}_dx{dx}_dy{dy}_h{dep}_n{nit000}-{nitend}_r{irecord}"
            "_dt{rdt}_cb{cbfr}_vi{visc}").format(**c)


def run_case(c, keep_dat=False):
    d = os.path.join(RUNS, case_name(c))
    if os.path.isdir(d):
        shutil.rmtree(d)
    os.makedirs(d)
    with open(os.path.join(d, "namelist"), "w") as f:
        f.write(NAMELIST.format(**c))
    env = dict(os.environ, OMP_NUM_THREADS="1")
    p = subprocess.run([REF_EXE], cwd=d, env=env, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("reference run failed for %s:\n%s\n%s" % (d, p.stdout, p.stderr))
    out = {}
    for line in p.stdout.splitlines():
        m = LINE.match(line)
        if m:
            out[m.group(1)] = m.group(2)
    if set(out) != {"ua", "va"}:
        raise RuntimeError("no checksum lines in output of %s:\n%s" % (d, p.stdout))
    dats = sorted(x for x in os.listdir(d) if x.startswith("go2d_") and x.endswith(".dat"))
    if not keep_dat:
        for x in dats:
            os.remove(os.path.join(d, x))
    return out, d, dats


def run_trace(c):
    """Per-step (istp, sum|sshn|, sum|un|, sum|vn|) of the trace build."""
    d = os.path.join(RUNS, "trace_" + case_name(c))
    if os.path.isdir(d):
        shutil.rmtree(d)
    os.makedirs(d)
    with open(os.path.join(d, "namelist"), "w") as f:
        f.write(NAMELIST.format(**c))
    env = dict(os.environ, OMP_NUM_THREADS="1")
    p = subprocess.run([TRACE_EXE], cwd=d, env=env, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("trace run failed for %s:\n%s\n%s" % (d, p.stdout, p.stderr))
    rows = []
    for line in p.stdout.splitlines():
        m = TRACE_LINE.match(line)
        if m:
            rows.append([float(m.group(1))] + [float(m.group(k)) for k in (2, 3, 4)])
    shutil.rmtree(d)
    if len(rows) != c["nitend"] - c["nit000"] + 1:
        raise RuntimeError("trace of %s has %d steps" % (case_name(c), len(rows)))
    return np.array(rows)


def trace_cases(big):
    cases = [base(cbfr=0.0015), base(cbfr=0.00015)]
    for (gi, gj) in [(3, 3), (7, 5), (12, 12), (36, 20)]:
        cases.append(base(jpiglo=gi, jpjglo=gj, nitend=400))
    cases.append(base(jpiglo=40, jpjglo=30, dx=750.0, dy=1300.0, dep=55.0, rdt=15.0, cbfr=0.004,
                      visc=30.0, nitend=300))
    cases.append(base(jpiglo=40, jpjglo=30, nit000=200, nitend=499))
    if big:
        cases.append(base(jpiglo=4098, jpjglo=4098, nitend=1000, irecord=1000000))
    return cases


def make_traces(big, jobs):
    path = os.path.join(HERE, "ref_step_checksums.npz")
    known = {}
    if os.path.exists(path):
        z = np.load(path)
        for k, nml in enumerate(json.loads(str(z["namelists"]))):
            known[case_name(nml)] = (nml, z["t%d" % k])
    todo = [c for c in trace_cases(big) if case_name(c) not in known]
    todo.sort(key=lambda c: c["jpiglo"] * c["jpjglo"] * (c["nitend"] - c["nit000"] + 1))

    def work(c):
        t = run_trace(c)
        print("trace", case_name(c), "%d steps" % len(t), flush=True)
        return c, t

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for c, t in ex.map(work, todo):
            known[case_name(c)] = (c, t)
            items = sorted(known.values(), key=lambda r: case_name(r[0]))
            arrays = {"t%d" % k: t for k, (_, t) in enumerate(items)}
            np.savez_compressed(path + ".tmp.npz", namelists=json.dumps([n for n, _ in items]),
                                source="oracle/_ref/nemolite2d_ref_trace built by oracle/ref.mk "
                                       "(flang -O2 -ffp-contract=off); columns istp, sum|sshn|, "
                                       "sum|un|, sum|vn| printed ES24.16", **arrays)
            os.replace(path + ".tmp.npz", path)


def base(**kw):
    c = dict(jpiglo=130, jpjglo=130, dx=1000.0, dy=1000.0, dep=100.0, nit000=1,
             nitend=2000, irecord=1000000, rdt=20.0, cbfr=0.0015, visc=0.1)
    c.update(kw)
    return c


def small_cases():
    cases = []
    steps128 = [1, 2, 3, 4, 5, 10, 20, 50, 100, 200, 300, 500, 750, 1000, 1500, 2000]
    for cbfr in (0.0015, 0.00015):
        for n in steps128:
            cases.append(base(nitend=n, cbfr=cbfr))
    # tiny and non-square grids (jpi, jpj = jpiglo-2, jpjglo-2), long enough
    # for reflections from every wall
    for (gi, gj) in [(3, 3), (4, 4), (5, 7), (8, 6), (12, 12), (34, 18), (18, 34)]:
        for n in (1, 2, 3, 7, 50, 400):
            cases.append(base(jpiglo=gi, jpjglo=gj, nitend=n))
    # other physics parameters: anisotropic spacing, depth, time step, larger
    # viscosity, a non-default start step (shifts the tide phase)
    for n in (1, 2, 5, 40, 300):
        cases.append(base(jpiglo=40, jpjglo=30, dx=750.0, dy=1300.0, dep=55.0,
                          rdt=15.0, cbfr=0.004, visc=30.0, nitend=n))
        cases.append(base(jpiglo=40, jpjglo=30, nit000=200, nitend=199 + n))
    return cases


def big_cases():
    return [base(jpiglo=4098, jpjglo=4098, nitend=n, irecord=1000000)
            for n in (1, 2, 10, 100, 1000)]


def field_cases():
    # snapshots every 500 steps at 128^2, every 10 steps on a small grid
    return [("128", base(nitend=2000, irecord=500)),
            ("40x30", base(jpiglo=40, jpjglo=30, nitend=60, irecord=10))]


def parse_dat(path):
    vals = np.loadtxt(path)
    return vals  # (jpi*jpj, 6), jj outer, ji inner


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--big", action="store_true", help="also run the 4096^2 cases")
    ap.add_argument("--jobs", type=int, default=6)
    ap.add_argument("--trace", action="store_true", help="only the per-step checksum traces")
    args = ap.parse_args()

    subprocess.run(["make", "-f", os.path.join(ROOT, "oracle", "ref.mk")], cwd=ROOT, check=True)
    os.makedirs(RUNS, exist_ok=True)
    if args.trace:
        return make_traces(args.big, args.jobs)

    jpath = os.path.join(HERE, "ref_checksums.json")
    known = {}
    if os.path.exists(jpath):
        with open(jpath) as f:
            known = {case_name(r["namelist"]): r for r in json.load(f)["runs"]}

    cases = small_cases() + (big_cases() if args.big else [])
    todo = [c for c in cases if case_name(c) not in known]

    def work(c):
        out, _, _ = run_case(c)
        print(case_name(c), out["ua"], out["va"], flush=True)
        return {"namelist": c, "ua": out["ua"], "va": out["va"]}

    def save():
        runs = sorted(known.values(), key=lambda r: (r["namelist"]["jpiglo"], r["namelist"]["jpjglo"],
                                                       case_name(r["namelist"])))
        with open(jpath + ".tmp", "w") as f:
            json.dump({"source": "oracle/_ref/nemolite2d_ref built by oracle/ref.mk "
                                 "(flang -O2 -ffp-contract=off) from original/nemolite2d.f90",
                       "format": "E16.8 strings as printed by original/nemolite2d.f90:82-85",
                       "runs": runs}, f, indent=1)
        os.replace(jpath + ".tmp", jpath)

    # smallest cases first; results are saved as each one finishes
    todo.sort(key=lambda c: (c["jpiglo"] * c["jpjglo"] * (c["nitend"] - c["nit000"] + 1)))
    with ThreadPoolExecutor(max_workers=args.jobs) as ex:
        for r in ex.map(work, todo):
            known[case_name(r["namelist"])] = r
            save()
    save()

    dat = os.path.join(HERE, "ref_go2d_10x10_00012.dat")
    if not os.path.exists(dat):
        c = base(jpiglo=12, jpjglo=12, nitend=12, irecord=12)
        _, d, dats = run_case(c, keep_dat=True)
        shutil.copyfile(os.path.join(d, "go2d_00012.dat"), dat)
        for x in dats:
            os.remove(os.path.join(d, x))
        print("wrote", dat, flush=True)

    for tag, c in field_cases():
        npz = os.path.join(HERE, "ref_fields_%s.npz" % tag)
        if os.path.exists(npz):
            continue
        _, d, dats = run_case(c, keep_dat=True)
        jpi, jpj = c["jpiglo"] - 2, c["jpjglo"] - 2
        steps, data = [], []
        for x in dats:
            v = parse_dat(os.path.join(d, x)).reshape(jpj, jpi, 6)
            if jpi > 64:   # decimate large grids: every 3rd column, every 3rd row
                v = v[::3, ::3, :]
            steps.append(int(x[5:10]))
            data.append(v[..., 3:6])   # ssh, u at T, v at T
            os.remove(os.path.join(d, x))
        stride = 3 if jpi > 64 else 1
        np.savez_compressed(npz, namelist=json.dumps(c), steps=np.array(steps),
                            stride=stride, fields=np.array(data))
        print("wrote", npz, flush=True)


if __name__ == "__main__":
    sys.exit(main())
