"""The constant-divisor division of the tiled HIP steppers (div_c in
meep-torch_amd/csrc/common.h) against IEEE float division, on the CPU.

The library passes the kernels y = RN(1/dx) only for an FP32 divisor whose
quotient sequence it has checked against IEEE division for every dividend of
magnitude in [2^-100, 2^100] and for zero (mt_div_recip_f32: all 1.68e9
positive bit patterns, threaded, cached per divisor), and 0 otherwise, which
makes the kernels divide; the kernels send any other dividend (tiny tails of
a wavefront, huge values) to true division.  Checked here: the library's
answer for the grid steps the configurations and the facade's usual
resolutions give, its refusal outside the divisor range, and --
independently -- a C program (tests/div_const_check.c, gcc, libm's FMA) that
tries every significand at 23 exponents across [2^-100, 2^100] and both
signs, and shows that the sequence does fail below 2^-100 (why the kernels
test the exponent).  The torch FP32 reference divides by dx with true
division, so bit-parity of the kernels rests on this."""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
RESOLUTIONS = (8, 10, 16, 20, 25, 30, 32, 40, 50, 64, 100)
DXS = [0.05, 0.025, 0.1] + [1.0 / r for r in RESOLUTIONS]


def test_library_verifies_every_usual_grid_step():
    from meep_torch_amd import _engine as E
    L = E.lib()
    for dx in DXS:
        b = np.float32(dx)
        y = L.mt_div_recip_f32(float(b))
        assert np.float32(y) == np.float32(1.0) / b, (dx, y)
    # outside [2^-16, 2^16] the exhaustive check would not cover every
    # dividend (subnormal intermediates): the library refuses, kernels divide
    assert L.mt_div_recip_f32(1.0e6) == 0.0
    assert L.mt_div_recip_f32(1.0e-7) == 0.0


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_markstein_constant_division_is_exact():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "div_const_check")
        subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, os.path.join(HERE, "div_const_check.c"), "-lm"],
                       check=True)
        res = subprocess.run([exe, "0.05", repr(1.0 / 30)], capture_output=True, text=True, timeout=600)
    print(res.stdout)
    assert res.returncode == 0, res.stdout
