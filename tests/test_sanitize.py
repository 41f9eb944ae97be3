"""The CPU oracle (the checker of every GPU test) built with
AddressSanitizer and UndefinedBehaviorSanitizer and run through a full
branch and bound with triangle, pentagonal and heptagonal cuts and a node
bound with fixings: no memory error and no undefined behaviour (every
report aborts the run)."""
import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORC = os.path.join(ROOT, "oracle")
SRCS = ["orc_ipm.c", "orc_cuts.c", "orc_bundle.c", "orc_heur.c", "orc_bound.c", "orc_many.c"]


def test_oracle_under_asan_and_ubsan():
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    tmp = tempfile.mkdtemp(prefix="orc_asan_")
    try:
        exe = os.path.join(tmp, "orc_asan")
        cmd = ["gcc", "-O1", "-g", "-std=c11", "-D_POSIX_C_SOURCE=200809L", "-fopenmp", "-fno-omit-frame-pointer",
               "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-I" + ORC,
               os.path.join(ROOT, "tests", "oracle_asan_main.c")] + [os.path.join(ORC, s) for s in SRCS] + \
              ["-o", exe, "-lm"]
        subprocess.run(cmd, check=True, capture_output=True)
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="halt_on_error=1")
        r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
        assert r.returncode == 0, r.stdout + r.stderr[-3000:]
        assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr
        assert "best" in r.stdout
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
