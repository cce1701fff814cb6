"""The kernels' shared-reciprocal division (Math<false>::divs, hush_amd/csrc/hush_device.h: one FMA correction
of a * RN(1/b)) is the correctly rounded quotient: the same IEEE operations in C (hardware FMA) against IEEE
division on 10^7 random operand pairs over wide exponent ranges, every quotient of small integers, and every
numerator significand for the divisors with the hard significands (all ones and its neighbours, 1, 1.5) and
16 random divisors (about 700 M quotients), and numerators from 2^-102 (the bound below which the residual can
leave the subnormal grid) over divisors of every normal exponent.  The special operands are checked on the
GPU (tests/test_gpu_math.py)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no C compiler")
def test_shared_reciprocal_division_is_correctly_rounded(tmp_path):
    exe = str(tmp_path / "divs_check")
    # -ffp-contract=off: the product a*y stays a separate rounding, as in the kernels
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-mfma", os.path.join(HERE, "divs_check.c"), "-o", exe, "-lm"],
                   check=True)
    r = subprocess.run([exe, "10000000", "16"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and r.stdout.strip() == "0", r.stdout + r.stderr
