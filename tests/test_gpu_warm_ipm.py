"""Warm-started bundle IPMs (bqb_ctx_set_ipm_warm): each bundle step's IPM
starts from the same node's previous solve instead of the reference's fixed
point (ipm_mc_pk.c:65-99).  The bounds stay valid and close to the cold
solves' (both stop at the IPM's 1e-2 gap, but the bundle trajectories then
part), fewer IPM iterations run, and every form proves the reference's
fixture optima."""
import os

import numpy as np
import pytest

from fixture_util import fixture_value
from parallel_biqbin_maxcut_amd import api, instances as I

pytestmark = pytest.mark.gpu


def _params(n0):
    p = api.make_params()
    p.TriIneq = 10 * n0                       # src/process_input.c:135-136
    return p


def _fixture(golden, name):
    path = os.path.join(golden, "rudy", name)
    return api.maxcut_L(api.read_rudy(path)), fixture_value(path + "-expected_output")


def _bounds(SP, xf, sx, warm, stream):
    ctx = api.BoundContext(SP, params=_params(SP.shape[0]), max_nodes=8 if stream else xf.shape[0])
    if warm:
        ctx.set_ipm_warm(True)
    seeds = np.arange(1, xf.shape[0] + 1, dtype=np.uint64)
    # root semantics: no give-up after the basic bound, so the bundle runs
    return ctx.bound(xf, sx, LB=0.0, is_root=1, seeds=seeds, stream=stream)


def test_root_bound_close_and_fewer_iterations(golden):
    SP, opt = _fixture(golden, "g05_60.0")
    n0 = SP.shape[0]
    xf = np.zeros((1, n0 - 1), np.int32)
    cold = _bounds(SP, xf, xf, False, False)
    warm = _bounds(SP, xf, xf, True, False)
    bc, bw = cold["bound"][0], warm["bound"][0]
    print(f"root bound cold {bc:.4f} warm {bw:.4f}; IPM iterations {cold['stats'][0][4]} -> {warm['stats'][0][4]}",
          flush=True)
    assert cold["status"][0] == 0 and warm["status"][0] == 0
    assert bw >= opt and abs(bw - bc) <= 5e-3 * bc
    assert warm["stats"][0][4] <= 0.85 * cold["stats"][0][4]


@pytest.mark.parametrize("stream", [False, True])
def test_nodes_of_different_sizes(stream):
    """16 nodes with 0-20 fixed variables; streamed through 8 slots, each
    slot takes nodes of other sizes, whose first bundle IPM must start cold."""
    adj = I.gnp_adj(50, 0.5, 7, weights="pm1")
    SP = api.maxcut_L(adj)
    n0 = SP.shape[0]
    rng = np.random.default_rng(5)
    xf = np.zeros((16, n0 - 1), np.int32)
    sx = np.zeros((16, n0 - 1), np.int32)
    for b in range(16):
        k = rng.integers(0, 21)
        idx = rng.choice(n0 - 1, size=k, replace=False)
        xf[b, idx] = 1
        sx[b, idx] = rng.integers(0, 2, size=k)
    cold = _bounds(SP, xf, sx, False, stream)
    warm = _bounds(SP, xf, sx, True, stream)
    assert np.all(cold["status"] == 0) and np.all(warm["status"] == 0)
    rel = np.abs(warm["bound"] - cold["bound"]) / np.maximum(1.0, np.abs(cold["bound"]))
    print(f"max relative bound difference {rel.max():.2e}; IPM iterations "
          f"{cold['stats'][:, 4].sum()} -> {warm['stats'][:, 4].sum()}", flush=True)
    assert rel.max() <= 1e-2
    assert warm["stats"][:, 4].sum() < cold["stats"][:, 4].sum()


@pytest.mark.parametrize("name,kw", [
    ("g05_60.0", {}),
    ("g05_80.0", {"batch": 16}),
    ("g05_80.0", {"batch": 64, "warm": True}),
    ("g05_60.0", {"warm": True, "slices": -1}),
    ("g05_60.0", {"slices": 4}),
])
def test_forms_prove_the_fixture(golden, name, kw):
    SP, opt = _fixture(golden, name)
    st, sol = api.solve(SP, _params(SP.shape[0]), seed=2020, ipm_warm=True, **kw)
    assert st.status == 0 and not st.stopped
    assert st.best == opt and api.evaluate_solution(SP, sol) == opt


def test_qubo_fixture(golden):
    import json
    base = os.path.join(golden, "qubos", "kcluster40_025_10_1.json")
    Q = I.read_sparse_json(base, "qubo")
    exp = json.load(open(base + "-expected_output.json"))
    SP = api.maxcut_L(I.qubo_to_maxcut_adj(Q))
    st, sol = api.solve(SP, _params(SP.shape[0]), seed=2020, batch=16, warm=True, ipm_warm=True)
    assert st.best == pytest.approx(exp["maxcut"]["computed_val"], abs=1e-6)
