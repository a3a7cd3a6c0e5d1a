"""The post-processing arithmetic of pfvs_amd/csrc/pfa_post.h (the fast
int64 expansion count, floor division, bounded dot products, Montgomery
products and signed reduction) and the int128 division of pfa_int.h (int64
operations when the operands fit) against their int128 / plain reference
forms, compiled for the host with g++ (tests/post_math_check.cpp)."""

import os
import shutil
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_post_math_fast_forms_equal_reference():
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    with tempfile.TemporaryDirectory() as td:
        exe = os.path.join(td, "post_math_check")
        subprocess.run([gxx, "-O2", "-std=c++17", "-ffp-contract=off", "-o", exe,
                        os.path.join(HERE, "post_math_check.cpp")], check=True)
        r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr
