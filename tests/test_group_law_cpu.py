"""The G1 group law's lazy forms on the CPU: the device headers built for the
host with g++ under UBSan (tests/native/group_law_test.cpp) against
reference formulas with normalized additions, on random and extreme inputs.
"""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT


def test_lazy_group_law_matches_reference(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "group_law_test")
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-fsanitize=undefined", "-fno-sanitize-recover=all",
                           "-I", os.path.join(ROOT, "tests", "native", "host"),
                           os.path.join(ROOT, "tests", "native", "group_law_test.cpp"), "-o", exe])
    out = subprocess.run([exe, "100000"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, (out.stdout, out.stderr[-3000:])
    assert out.stdout.split() == ["0", "0", "0"]
