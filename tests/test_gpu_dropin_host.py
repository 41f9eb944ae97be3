"""The drop-in boundary from a C host linked against libbiqbin_gpu.so, as
the reference's B&B links the bounding files (tests/dropin_host.c): the
host owns Bab_LBGet/updateSolution, params, diff and the global cut lists,
and calls Evaluate (rank 0 and 1), createSubproblem and fct_eval by their
reference signatures.  A second build defines the reference's heuristic
override (wrapped_heuristic), which must replace the device GW.

Checked: the root bound of g05_60.0 against the oracle (2e-4 relative, the
trajectory's own sensitivity with pentagonal/heptagonal cuts), the
incumbent raised through the host's updateSolution to the fixture optimum,
the root gap written to the host's diff, the children's bounds valid, and
fct_eval with non-empty Cuts/Pent_Cuts/Hepta_Cuts against the oracle's
fct_eval within 1e-6."""
import json
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

import oracle_py as O
from fixture_util import fixture_value
from parallel_biqbin_maxcut_amd import api

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "parallel_biqbin_maxcut_amd")


def _build(tmp, override):
    exe = os.path.join(tmp, "host_ov" if override else "host")
    cmd = ["gcc", "-O1", "-std=c11", "-I" + os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "dropin_host.c"),
           "-o", exe, "-L" + PKG, "-lbiqbin_gpu", "-Wl,-rpath," + PKG]
    if override:
        cmd.insert(1, "-DOVERRIDE")
    subprocess.run(cmd, check=True, capture_output=True)
    return exe


def _run(exe, graph):
    r = subprocess.run([exe, graph], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return {d["step"]: d for d in (json.loads(l) for l in r.stdout.splitlines() if l.startswith("{"))}


@pytest.fixture(scope="module")
def hosts():
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    tmp = tempfile.mkdtemp(prefix="bqb_host_")
    yield _build(tmp, False), _build(tmp, True)
    shutil.rmtree(tmp, ignore_errors=True)


def test_host_linked_evaluate_and_fct_eval(hosts, golden):
    graph = os.path.join(golden, "rudy", "g05_60.0")
    out = _run(hosts[0], graph)
    SP = api.maxcut_L(api.read_rudy(graph))
    n0 = SP.shape[0]
    opt = fixture_value(graph + "-expected_output")
    root = out["root"]
    print(json.dumps(out), flush=True)
    P = O.default_params(TriIneq=10 * n0)
    z = np.zeros(n0 - 1, np.int32)
    o, _, _ = O.sdpbound(SP, z, z, 0.0, 0.0, 1, int(O.key(2020, 0)), P)
    assert abs(root["bound"] - o.bound) <= 2e-4 * o.bound, (root["bound"], o.bound)
    assert root["LB"] == opt and root["updates"] >= 1
    assert root["diff"] > 0.0 and root["n"] == n0
    assert max(out["child0"]["bound"], out["child1"]["bound"]) >= opt
    # fct_eval on the root subproblem with the host's cuts
    fe = out["fct_eval"]
    L, _ = O.create_subproblem(SP, z, z)
    n = L.shape[0]
    cuts = O.cut_arrays(tris=[(2, 1, 0, 1), (5, 3, 1, 2), (n - 1, 4, 2, 4)],
                        pents=[(2, [3 * a + 1 for a in range(5)])], hepts=[(3, [2 * a for a in range(7)])])
    f, _, g, _ = O.fct_eval(L, cuts, np.array([0.125, 0.5, 0.0625, 0.25, 0.2]))
    assert abs(fe["f"] - f) <= 1e-6 * abs(f), (fe["f"], f)
    assert np.allclose(fe["g"], g, rtol=0, atol=1e-6)


def test_host_heuristic_override_replaces_device_gw(hosts, golden):
    graph = os.path.join(golden, "rudy", "g05_60.0")
    out = _run(hosts[1], graph)
    SP = api.maxcut_L(api.read_rudy(graph))
    ones = np.ones(SP.shape[0] - 1, np.int32)
    v = api.evaluate_solution(SP, ones)
    root = out["root"]
    print(json.dumps(root), flush=True)
    assert root["heur_calls"] >= 1
    assert root["LB"] == max(0.0, v)          # only the override's cut reached the incumbent
    assert root["bound"] >= fixture_value(graph + "-expected_output")
