"""The decision tape's word layout and the Reflect case split
(0buf_amd/csrc/k_tape.h), host/device code built here with g++: every field
round-trips every value it can hold without touching the others, an
out-of-range value marks the word unreplayable, and the recorded Reflect
case applied to the position reproduces the push's own Reflect bit for bit
(signed zeros and the exact 0 and L edges included)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_tape_word_fields_and_reflect_cases(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = tmp_path / "tape_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-ffp-contract=off", "-I",
                    os.path.join(ROOT, "0buf_amd", "csrc"),
                    os.path.join(HERE, "cpu", "tape_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], stdout=subprocess.PIPE, text=True)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout
