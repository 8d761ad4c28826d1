"""flux3's divisions by 12 and 60 run on the device as a multiply by the
rounded reciprocal plus one FMA correction, with the IEEE division kept for
tiny, infinite and NaN inputs (kernels_std.hip, div_const).  This checks, over
all 2^32 FP32 inputs, that the result is the IEEE quotient bit for bit (a C
restatement of the device function, compiled here); the GPU tests check the
kernels against the oracle, with tiny and subnormal values among them."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    exe = str(tmp_path_factory.mktemp("divc") / "div_const_check")
    flags = ["-O2", "-fno-fast-math", "-ffp-contract=off", "-pthread"]
    probe = subprocess.run([cc, "-mfma", "-E", "-x", "c", os.devnull], capture_output=True)
    if probe.returncode == 0:
        flags.append("-mfma")
    subprocess.run([cc, *flags, os.path.join(HERE, "div_const_check.c"), "-o", exe, "-lm"],
                   check=True)
    return exe


@pytest.mark.parametrize("c", ["12", "60"])
def test_division_by_constant_is_the_ieee_quotient_for_every_float(checker, c):
    r = subprocess.run([checker, c], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mismatches=0" in r.stdout
    fast = int(r.stdout.split("fast=")[1].split()[0])
    assert fast > 4_000_000_000  # the three-operation path covers nearly every input
