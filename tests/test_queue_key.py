"""The arrival kernels trust a departure record only after the checks of
0buf_amd/csrc/k_queue_key.h (halo index inside the tile's box, a claim made
for that halo cell, rank below the tile's count for it, slot not below the
cell's bucket-held slots).  The header is host/device code; this test builds
it on the CPU with g++ and checks the key round trip, that the slots of a
tile's valid records fill each destination's claimed run exactly once, and
that damaged records are refused.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_queue_key_and_arrival_guard(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = tmp_path / "queue_key_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "0buf_amd", "csrc"),
                    os.path.join(HERE, "cpu", "queue_key_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], stdout=subprocess.PIPE, text=True)
    assert out.returncode == 0 and out.stdout.startswith("ok"), out.stdout
    assert int(out.stdout.split()[1]) > 10_000
