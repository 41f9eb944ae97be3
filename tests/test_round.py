"""The directed-rounding arithmetic of the certified bounds (bqb_round.h)
gives exactly the results of the host's hardware rounding modes (upward and
downward) on two million random operand pairs and on edge cases (zeros,
subnormals, neighbours of 1); products and quotients below 2^-960 may be one
step wider, on the correct side.  CPU only: the header is compiled with gcc
at -O0 (gcc's optimiser moves floating-point operations across fesetround).
"""
import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_directed_rounding_matches_hardware_modes():
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    tmp = tempfile.mkdtemp(prefix="bqb_round_")
    try:
        exe = os.path.join(tmp, "round_check")
        subprocess.run(["gcc", "-O0", "-std=c11", "-frounding-math", "-ffp-contract=off",
                        "-I" + os.path.join(ROOT, "parallel_biqbin_maxcut_amd", "csrc"),
                        os.path.join(ROOT, "tests", "round_check.c"), "-o", exe, "-lm"],
                       check=True, capture_output=True)
        r = subprocess.run([exe], capture_output=True, text=True)
        assert r.returncode == 0 and r.stdout.strip() == "0", r.stdout + r.stderr
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
