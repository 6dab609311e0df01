"""The constant-divisor division of the HIP steppers (div_c in
meep-torch_amd/csrc/common.h) against IEEE float division, on the CPU: a C
program (tests/div_const_check.c, compiled here with gcc and the FMA of
libm) tries every significand at 23 exponents across [2^-100, 2^100], both
signs, for the grid steps the configs use.  The torch FP32 reference divides
by dx with true division, so bit-parity of the tiled kernels rests on this."""
import os
import shutil
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_markstein_constant_division_is_exact():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "div_const_check")
        subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, os.path.join(HERE, "div_const_check.c"), "-lm"],
                       check=True)
        res = subprocess.run([exe, "0.05", "0.025", "0.1"], capture_output=True, text=True, timeout=600)
    print(res.stdout)
    assert res.returncode == 0, res.stdout
