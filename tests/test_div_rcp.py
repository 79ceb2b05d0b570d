"""The device's exact division by a precomputed reciprocal (div_rcp, and
div_rcp_acc wherever its gathered guard holds, in
nemolite2d-cuda_amd/csrc/nemo2d_core.h), restated on the host in
tests/native/div_rcp_check.c, gives the bits of the IEEE quotient x / d on
random, near-midpoint and special operands; so does the one-correction form
(div_rcp1_acc) on divisors that pass the library's test for it
(|d*RN(1/d) - 1| <= 2^-54, which makes RN(x*y) a faithful quotient).  The reference divides with
IEEE division throughout (original/nemolite2d.f90:403, :505, :597, :737);
the device kernels that use the reciprocal form must give the same bits."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_division_by_reciprocal_is_exact(tmp_path):
    exe = str(tmp_path / "div_rcp_check")
    subprocess.check_call(["gcc", "-O2", "-ffp-contract=off", "-std=c11", "-o", exe,
                           os.path.join(HERE, "native", "div_rcp_check.c"), "-lm"])
    for seed in (1, 2, 3):
        out = subprocess.check_output([exe, "400000", str(seed)], text=True).split()
        tried, bad, guarded, tried1 = (int(v) for v in out[:4])
        assert tried > 2_000_000 and guarded > tried // 2 and tried1 > tried // 3
        assert bad == 0, "%d of %d quotients differ from IEEE division" % (bad, tried)
