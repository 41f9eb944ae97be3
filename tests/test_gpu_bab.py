"""End to end on the GPU: the single-process node queue (bqb_solve) and the
drop-in Evaluate() prove the same optimum as the reference's own fixtures
(tests/rudy, tests/qubos of the reference) and as brute force, bit-exact."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle_py as O
from fixture_util import brute_maxcut, fixture_root_bound, fixture_solution, fixture_value
from parallel_biqbin_maxcut_amd import api, instances as I

pytestmark = pytest.mark.gpu

PKG = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "parallel_biqbin_maxcut_amd")


def reference_params(n0, **kw):
    p = api.make_params(**kw)
    if p.adjust_TriIneq:
        p.TriIneq = 10 * n0                  # src/process_input.c:135-136
    return p


@pytest.mark.parametrize("seed,batch", [(1, 1), (2, 8), (3, 32)])
def test_small_graphs_match_brute_force(seed, batch):
    adj = I.gnp_adj(16, 0.5, seed, weights="int10")
    SP = api.maxcut_L(adj)
    st, sol = api.solve(SP, reference_params(16), seed=seed, batch=batch)
    assert st.status == 0 and not st.stopped
    assert st.best == brute_maxcut(adj)
    assert api.evaluate_solution(SP, sol) == st.best


def _rudy(golden, name, batch):
    path = os.path.join(golden, "rudy", name)
    exp = path + "-expected_output"
    SP = api.maxcut_L(api.read_rudy(path))
    st, sol = api.solve(SP, reference_params(SP.shape[0]), seed=2020, batch=batch)
    assert st.status == 0
    assert st.best == fixture_value(exp)
    assert api.evaluate_solution(SP, sol) == st.best
    fsol = np.zeros(SP.shape[0] - 1, dtype=np.int32)
    fsol[np.array(fixture_solution(exp)) - 1] = 1
    assert api.evaluate_solution(SP, fsol) == st.best
    return st, sol, exp


# The full root bound with pentagonal/heptagonal cuts moves by up to 2e-4
# relative under a 1e-15 relative change of its input (the oracle's own
# spread on g05_60.0), so the printed root bound is compared at that level;
# the optimum and the node count are exact.
ROOT_BOUND_REL = 2e-4


def test_fixture_g05_60_root_only(golden):
    st, sol, exp = _rudy(golden, "g05_60.0", 16)
    assert st.nodes == 1                                  # "Nodes = 1"
    ref = fixture_root_bound(exp)
    assert abs(st.root_bound - ref) <= ROOT_BOUND_REL * ref + 0.005


@pytest.mark.parametrize("name,batch", [("g05_80.0", 16), ("g05_100.4", 64)])
def test_fixture_rudy_optimum(golden, name, batch):
    _rudy(golden, name, batch)


def test_fixture_kcluster40_qubo(golden):
    base = os.path.join(golden, "qubos", "kcluster40_025_10_1.json")
    Q = I.read_sparse_json(base, "qubo")
    exp = json.load(open(base + "-expected_output.json"))
    SP = api.maxcut_L(I.qubo_to_maxcut_adj(Q))
    st, sol = api.solve(SP, reference_params(SP.shape[0]), seed=2020, batch=16)
    assert st.status == 0
    assert st.best == pytest.approx(exp["maxcut"]["computed_val"], abs=1e-6)
    y = I.maxcut_solution_to_qubo([i + 1 for i in range(len(sol)) if sol[i]], Q.shape[0])
    assert y @ Q @ y == pytest.approx(exp["qubo"]["computed_val"], abs=1e-6)


def test_batch_size_does_not_change_the_optimum():
    adj = I.gnp_adj(40, 0.5, 11, weights="pm1")
    SP = api.maxcut_L(adj)
    bests = {b: api.solve(SP, reference_params(40), seed=3, batch=b)[0].best for b in (1, 4, 64)}
    assert len(set(bests.values())) == 1, bests
    out, _ = O.bab(SP, O.reference_params_file(TriIneq=400), seed=3)
    assert out.best == bests[1]


def test_cli_reports_the_fixture(golden, tmp_path):
    exe = os.path.join(PKG, "biqbin_gpu")
    r = subprocess.run([exe, os.path.join(golden, "rudy", "g05_60.0"), os.path.join(golden, "params"), "8"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Maximum value = 536" in r.stdout
    rb = float(r.stdout.split("Root node bound = ")[1].split()[0])
    assert abs(rb - 536.95) <= ROOT_BOUND_REL * 536.95 + 0.005


@pytest.mark.parametrize("opts", [["-w", "-p", "auto"], ["-w", "-p", "4", "-g"], ["-c"], ["-w", "-i"]])
def test_cli_options_prove_the_fixture(golden, opts):
    """The CLI's solver forms (inherited cuts, automatic or bf16 precision
    with certified bounds, graphs, warm-started IPMs) prove the same optimum."""
    exe = os.path.join(PKG, "biqbin_gpu")
    r = subprocess.run([exe] + opts + [os.path.join(golden, "rudy", "g05_60.0"), os.path.join(golden, "params")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Maximum value = 536" in r.stdout
    bad = subprocess.run([exe, "-p", "9", "x", "y"], capture_output=True, text=True, timeout=60)
    assert bad.returncode == 1 and "Usage" in bad.stderr


def test_dropin_evaluate_root_node(golden):
    """Evaluate() with the reference's signature on the root node: same
    bound as the batch entry point with the drop-in's stream key."""
    SP = api.maxcut_L(api.read_rudy(os.path.join(golden, "rudy", "g05_60.0")))
    n0 = SP.shape[0]
    L = api.N.lib()
    L.bqb_dropin_release()
    PPbuf = np.zeros(n0 * n0)
    sp = api.Problem(SP.ctypes.data_as(api.N.dptr), n0, 0, 0, 0, 0)
    pp = api.Problem(PPbuf.ctypes.data_as(api.N.dptr), n0, 0, 0, 0, 0)
    node = api.BabNode()
    bound = L.Evaluate(node, sp, pp, 0)
    sol = np.zeros(n0 - 1, dtype=np.int32)
    lb = L.bqb_dropin_incumbent(api._i(sol))
    L.bqb_dropin_release()
    assert pp.n == n0
    ctx = api.BoundContext(SP, params=api.default_params(), max_nodes=1)
    xf = np.zeros((1, n0 - 1), np.int32)
    r = ctx.bound(xf, xf, LB=0.0, is_root=1, seeds=O.key(2020, 0))
    assert bound == r["bound"][0]
    assert lb == r["LB"][0] and lb == api.evaluate_solution(SP, sol)
    assert np.array_equal(np.array(node.fracsol[: n0 - 1]), r["fracsol"][0])


@pytest.mark.parametrize("name", ["g05_60.0", "g05_80.0"])
def test_fixture_rudy_json_copies(golden, name):
    """The reference's adjacency-JSON copies of the rudy fixtures (read as
    biqbin_base.py:33-45, 76-89 does) prove the same optimum."""
    path = os.path.join(golden, "rudy", name)
    adj = I.read_sparse_json(path + ".json", "adjacency")
    SP = api.maxcut_L(adj)
    st, sol = api.solve(SP, reference_params(SP.shape[0]), seed=2020, batch=32)
    assert st.status == 0
    assert st.best == fixture_value(path + "-expected_output")


def test_fixture_kcluster80_qubo(golden):
    """kcluster80 (QUBO on 80 variables -> Max-Cut on 81 vertices): the
    fixture's optimum 16400 / QUBO -16400 and a solution attaining it."""
    base = os.path.join(golden, "qubos", "kcluster80_025_20_1.json")
    Q = I.read_sparse_json(base, "qubo")
    exp = json.load(open(base + "-expected_output.json"))
    SP = api.maxcut_L(I.qubo_to_maxcut_adj(Q))
    st, sol = api.solve(SP, reference_params(SP.shape[0]), seed=2020, batch=64)
    print(f"kcluster80: {st.nodes} nodes, {st.seconds:.2f} s", flush=True)
    assert st.status == 0
    assert st.best == pytest.approx(exp["maxcut"]["computed_val"], abs=1e-6)
    y = I.maxcut_solution_to_qubo([i + 1 for i in range(len(sol)) if sol[i]], Q.shape[0])
    assert y @ Q @ y == pytest.approx(exp["qubo"]["computed_val"], abs=1e-6)
