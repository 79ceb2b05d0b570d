"""Comparison tools for the model's outputs (host only).

    python -m nemo2d.compare dat REFERENCE SAMPLE [-e EPS] [-o ROWS.csv] [-f]
    python -m nemo2d.compare checksums LEFT RIGHT [-e EPS] [-r RTOL]

``dat`` compares two go2d snapshot files (x, y, depth, ssh, u, v per T
point, original/nemolite2d.f90:773-799) the way the reference's
tools/diff.py does (tools/diff.py:91-109): a row is wrong when |ssh|, |u| or
|v| differs by more than EPS; prints ``Success`` or the count and the summed
absolute differences, and can store the wrong rows.  ``checksums`` compares
two checksum tables: CSVs with ``ua_checksum``/``va_checksum`` columns (the
reference's benchmark extractor output, matched on ``scale`` when both have
it) or the per-step files of this package's Fortran driver (istp, sum|ssh|,
sum|u|, sum|v|).  Unlike tools/csv-checksum.py:82-83, which subtracts the
left column from itself, left is compared with right.  Exit status 0 when
everything agrees, 1 otherwise.
"""
import argparse
import os
import sys

import numpy as np

DAT_COLUMNS = ("x", "y", "depth", "ssh", "u_velocity", "v_velocity")


def read_dat(path):
    """(n, 6) array of a go2d file (its empty records skipped)."""
    import pandas as pd
    df = pd.read_csv(path, header=None, sep=r"\s+", skip_blank_lines=True, dtype=np.float64,
                     engine="c")
    a = df.to_numpy()
    if a.ndim != 2 or a.shape[1] != 6:
        raise ValueError("%s: not a 6-column snapshot file" % path)
    return a


def compare_dat(ref, sample, eps=0.0):
    """(wrong-row mask, summed |delta| of ssh, u, v) for two (n, 6) arrays."""
    if ref.shape != sample.shape:
        raise ValueError("row counts differ: %d vs %d" % (len(ref), len(sample)))
    d = np.abs(ref - sample)
    wrong = (d[:, 3] > eps) | (d[:, 4] > eps) | (d[:, 5] > eps)
    return wrong, d[:, 3:6].sum(axis=0)


def read_checksums(path):
    """{key: (ua, va)} from an extractor CSV (key: scale, else row index) or
    {istp: (sum|ssh|, sum|u|, sum|v|)} from a driver per-step file."""
    with open(path) as f:
        head = f.readline()
    if "ua_checksum" in head:
        import pandas as pd
        df = pd.read_csv(path)
        keys = df["scale"].tolist() if "scale" in df.columns else list(range(len(df)))
        return {k: (float(a), float(b)) for k, a, b in zip(keys, df["ua_checksum"], df["va_checksum"])}
    a = np.loadtxt(path, comments="#", ndmin=2)
    return {int(r[0]): tuple(float(x) for x in r[1:]) for r in a}


def compare_checksums(left, right, eps=0.0, rtol=0.0):
    """Rows (key, left, right) where any value differs by more than
    eps + rtol*|right|, and the keys present on one side only."""
    bad = []
    for k in sorted(set(left) & set(right), key=str):
        a, b = np.array(left[k]), np.array(right[k])
        if a.shape != b.shape or np.any(np.abs(a - b) > eps + rtol * np.abs(b)):
            bad.append((k, left[k], right[k]))
    missing = sorted(set(left) ^ set(right), key=str)
    return bad, missing


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m nemo2d.compare", description=__doc__.split("\n\n")[0])
    sub = ap.add_subparsers(dest="what", required=True)
    d = sub.add_parser("dat", help="compare two go2d snapshot files")
    d.add_argument("reference")
    d.add_argument("sample")
    d.add_argument("-e", "--epsilon", type=float, default=0.0)
    d.add_argument("-o", "--output", help="write the wrong rows (both sides) as CSV")
    d.add_argument("-f", "--force", action="store_true", help="overwrite --output")
    c = sub.add_parser("checksums", help="compare two checksum tables")
    c.add_argument("left")
    c.add_argument("right")
    c.add_argument("-e", "--epsilon", type=float, default=0.0)
    c.add_argument("-r", "--rtol", type=float, default=0.0)
    args = ap.parse_args(argv)

    if args.what == "dat":
        for p in (args.reference, args.sample):
            if not os.path.isfile(p):
                print("Error: %s is not a file" % p)
                return 1
        if args.output and os.path.exists(args.output) and not args.force:
            print("Error: output file %s already exists. Please use -f/--force" % args.output)
            return 1
        ref, smp = read_dat(args.reference), read_dat(args.sample)
        wrong, sums = compare_dat(ref, smp, args.epsilon)
        n = int(wrong.sum())
        if n == 0:
            print("Success")
            return 0
        print("Failure. %d incorrect rows." % n)
        print("Summary:\n\tssh = %r\n\tu_velocity = %r\n\tv_velocity = %r" % tuple(float(x) for x in sums))
        if args.output:
            idx = np.nonzero(wrong)[0]
            with open(args.output, "w") as f:
                f.write("row,x,y,ssh_a,ssh_b,u_velocity_a,u_velocity_b,v_velocity_a,v_velocity_b\n")
                for k in idx:
                    f.write("%d,%r,%r,%r,%r,%r,%r,%r,%r\n" % (k, ref[k, 0], ref[k, 1], ref[k, 3], smp[k, 3],
                                                              ref[k, 4], smp[k, 4], ref[k, 5], smp[k, 5]))
            print("See '%s' for details." % args.output)
        return 1

    left, right = read_checksums(args.left), read_checksums(args.right)
    bad, missing = compare_checksums(left, right, args.epsilon, args.rtol)
    print("Comparing checksums using epsilon %r, rtol %r" % (args.epsilon, args.rtol))
    if missing:
        print("%d rows on one side only: %s" % (len(missing), missing[:10]))
    if bad:
        print("%d incorrect rows" % len(bad))
        for k, a, b in bad[:20]:
            print("  %s: left %s  right %s" % (k, a, b))
    if not bad and not missing:
        print("Correct!")
        return 0
    return 1


if __name__ == "__main__":
    sys.exit(main())
