"""The FP64 push divides by run constants (h per axis, tau h h, 12) with
divc (0buf_amd/csrc/k_push_math.h): q0 = x r, r = RN(1/d), corrected by one
FMA with the exactly formed remainder x - q0 d.  Markstein's theorem gives
RN(x/d) when q0 is a faithful quotient; q0 = RN(x RN(1/d)) is not faithful
for every x and d in principle, so the bit-identity with IEEE division that
the gather indices, weights and crossing codes rely on is checked here on
samples: the divisors of every shipped workload (h = 0.1 and the products
tau h h), near-multiples of them, and random divisors.  The probe is C built
with -ffp-contract=off, the same sequence the device code uses.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpu", "divc_check.c")


def _run(tmp_path, extra, n):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    exe = tmp_path / ("divc" + "".join(e.strip("-") for e in extra))
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", *extra, SRC, "-lm", "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe), str(n)], stdout=subprocess.PIPE, text=True)
    bad, total = map(int, out.stdout.split())
    return bad, total


def test_divc_equals_ieee_division(tmp_path):
    bad, total = _run(tmp_path, [], 3_000_000)
    assert total == 84_000_000
    assert bad == 0, f"{bad} of {total} quotients differ from x/d"


def test_divc_probe_sees_uncorrected_quotients(tmp_path):
    """planted mistake: without the FMA correction the probe must report
    mismatches, so a pass above is not vacuous"""
    bad, total = _run(tmp_path, ["-DUNCORRECTED"], 100_000)
    assert bad > 0.01 * total
