"""The fast quotient of the corrective kernel (hicar_amd/csrc/fdiv.hpp: q0 =
x*y, rn = q0*d - x, q = q0 - rn*y, y the hardware reciprocal refined by one
Newton step) against the IEEE division on the device, for all 2^46 pairs of
significands (x, d in [1, 2)), with the reciprocal's exact scaling over
d's exponents [-60, 40] and a negative numerator per divisor.  The kernels
take the fast quotient only where every one of its operations scales exactly
with the operands' exponents, so this covers every quotient they take that
way."""
import ctypes
import time

import pytest

from hicar_amd import _lib

pytestmark = pytest.mark.gpu

CHUNK = 1 << 16


def test_fast_quotient_is_the_ieee_quotient_for_every_significand_pair():
    lib = _lib.load()
    bad, first = ctypes.c_ulonglong(), ctypes.c_ulonglong()
    total = 0
    t0 = time.perf_counter()
    for n, d0 in enumerate(range(0, 1 << 23, CHUNK)):
        rc = lib.icar_adv_verify_division(d0, CHUNK, ctypes.byref(bad), ctypes.byref(first))
        assert rc == 0, lib.icar_adv_last_error(None)
        total += bad.value
        assert bad.value == 0, f"divisors from {d0:#x}: {bad.value} mismatches, first " \
                               f"{first.value:#018x}"
        if n % 16 == 15:
            print(f"divisors < {d0 + CHUNK:#08x}: no mismatch ({time.perf_counter() - t0:.0f} s)",
                  flush=True)
    assert total == 0


def test_fused_kernels_with_the_fast_quotient_match_the_oracle():
    """The opt-in fast quotient (ICAR_ADV_FAST_DIV=1, read once per process)
    in a process of its own: tests/fastdiv_check.py runs the fused kernels on
    random domains, extreme and negative species (the fallback), the crop of
    the 10-species domain and 16-bit storage, each bit for bit against the
    oracle."""
    import os
    import subprocess
    import sys
    here = os.path.dirname(os.path.abspath(__file__))
    env = dict(os.environ, ICAR_ADV_FAST_DIV="1")
    r = subprocess.run([sys.executable, os.path.join(here, "fastdiv_check.py")], env=env,
                       capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("bit-identical") == 7
