"""Snapshot files (go2d_NNNNN.dat) and the comparison tools (CPU)."""
import os

import numpy as np

import nemo2d
from oracle import Oracle

HERE = os.path.dirname(os.path.abspath(__file__))


def test_snapshot_text_equals_reference_file():
    """The library's formatter on the oracle's fields after 12 steps of the
    10x10 basin gives the reference binary's go2d_00012.dat byte for byte
    (E20.7E4 fields, u and v averaged to T points, an empty record per row)."""
    o = Oracle(10, 10)
    o.run(1, 12)
    txt = nemo2d.format_go2d(o.arr("xt"), o.arr("yt"), o.arr("ht"), o.arr("sshn"), o.arr("un"),
                             o.arr("vn"))
    with open(os.path.join(HERE, "golden", "ref_go2d_10x10_00012.dat"), "rb") as f:
        ref = f.read()
    assert txt == ref


def test_snapshot_number_format_edge_cases():
    """Signed zero, rounding carry, ties to even, large exponents: the forms
    flang's (6e20.7e4) gives for these values."""
    vals = [0.0, -0.0, 1.0, -2.5e-300, 1.0e300, 9.99999995, 0.5, 123456785.0, 12345665.0,
            12345675.0]
    expect = ["0.0000000E+0000", "-0.0000000E+0000", "0.1000000E+0001", "-0.2500000E-0299",
              "0.1000000E+0301", "0.1000000E+0002", "0.5000000E+0000", "0.1234568E+0009",
              "0.1234566E+0008", "0.1234568E+0008"]
    # one-row files with the value in the ssh column
    for v, e in zip(vals, expect):
        one = np.zeros((1, 1))
        txt = nemo2d.format_go2d(one, one, one, np.full((1, 1), v), np.zeros((1, 2)), np.zeros((2, 1)))
        line = txt.split(b"\n")[0]
        assert len(line) == 120
        assert line[60:80].decode() == e.rjust(20), (v, line[60:80])


def _dat(o):
    return nemo2d.format_go2d(o.arr("xt"), o.arr("yt"), o.arr("ht"), o.arr("sshn"), o.arr("un"),
                              o.arr("vn"))


def test_compare_dat_finds_changed_rows(tmp_path, capsys):
    from nemo2d import compare
    o = Oracle(20, 15)
    o.run(1, 30)
    a = tmp_path / "a.dat"
    a.write_bytes(_dat(o))
    assert compare.main(["dat", str(a), str(a)]) == 0
    assert "Success" in capsys.readouterr().out
    o.arr("sshn")[3, 4] += 1e-3      # one T point
    o.arr("vn")[10, 2] += 5e-4       # moves the v average of two points
    b = tmp_path / "b.dat"
    b.write_bytes(_dat(o))
    out = tmp_path / "rows.csv"
    assert compare.main(["dat", str(a), str(b), "-o", str(out)]) == 1
    assert "Failure. 3 incorrect rows." in capsys.readouterr().out
    assert len(out.read_text().splitlines()) == 4
    # within epsilon: accepted
    assert compare.main(["dat", str(a), str(b), "-e", "1e-2"]) == 0


def test_compare_checksums_left_against_right(tmp_path, capsys):
    from nemo2d import compare
    l, r = tmp_path / "l.csv", tmp_path / "r.csv"
    l.write_text("scale,ua_checksum,va_checksum\n128,0.1,2.0\n256,0.3,4.0\n")
    r.write_text("scale,ua_checksum,va_checksum\n128,0.1,2.0\n256,0.3,4.5\n")
    assert compare.main(["checksums", str(l), str(l)]) == 0
    assert compare.main(["checksums", str(l), str(r)]) == 1
    assert "1 incorrect rows" in capsys.readouterr().out
    assert compare.main(["checksums", str(l), str(r), "-e", "0.6"]) == 0
    # per-step files of the Fortran driver
    p1, p2 = tmp_path / "a.txt", tmp_path / "b.txt"
    p1.write_text("# istp ssh u v\n1 1.0 2.0 3.0\n2 1.5 2.5 3.5\n")
    p2.write_text("# istp ssh u v\n1 1.0 2.0 3.0\n2 1.5 2.5 3.5000000000001\n")
    assert compare.main(["checksums", str(p1), str(p2)]) == 1
    assert compare.main(["checksums", str(p1), str(p2), "-r", "1e-12"]) == 0
