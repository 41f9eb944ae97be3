"""Launch trains captured as HIP graphs (bqb_ctx_set_graphs): the same
kernels with the same arguments in the same order as the eager launches, so
every result must be bitwise equal.  Each graph is replayed many times with
eager work and host synchronisations in between, and every replay is checked
against the eager run."""
import numpy as np
import pytest

from parallel_biqbin_maxcut_amd import api, instances as I

pytestmark = pytest.mark.gpu

KEYS = ("bound", "fracsol", "xbest", "LB", "improved", "diff", "status", "stats")


def _workload(n0, p, weights, nodes, seed):
    SP = api.maxcut_L(I.gnp_adj(n0, p, seed, weights=weights))
    xf, sx = I.random_fixings(n0, min(40, n0 // 3), nodes, seed + 1)
    seeds = np.arange(1, nodes + 1, dtype=np.uint64) + np.uint64(1000 * seed)
    return SP, xf, sx, seeds


def _same(a, b):
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("slots", [1, 16])
def test_graph_trains_bitwise_equal_to_eager(slots):
    SP, xf, sx, seeds = _workload(60, 0.5, "pm1", 48, 3)
    P = api.make_params()
    eager = api.BoundContext(SP, params=P, max_nodes=slots)
    ref = eager.bound(xf, sx, LB=20.0, diff=5.0, seeds=seeds, stream=True)
    c0 = eager.counters()
    assert c0["graph_launches"] == 0 and c0["eager_phases"] > 0
    g = api.BoundContext(SP, params=P, max_nodes=slots)
    g.set_graphs(True)
    for rep in range(3):
        out = g.bound(xf, sx, LB=20.0, diff=5.0, seeds=seeds, stream=True)
        _same(out, ref)
        # eager work on the other context between the replays
        again = eager.bound(xf[:8], sx[:8], LB=20.0, diff=5.0, seeds=seeds[:8], stream=True)
        for k in KEYS:
            assert np.array_equal(again[k], ref[k][:8]), k
    c = g.counters()
    print(slots, c, flush=True)
    assert c["graph_launches"] > 0 and c["graph_phases"] > 10 * c["graph_launches"] // 2
    assert c["eager_phases"] <= c["advances"]          # only the PH_LOAD of each upload
    assert c["graph_instantiations"] < c["graph_launches"]
    eager.close()
    g.close()


def test_graphs_with_mixed_precision_and_model():
    from parallel_biqbin_maxcut_amd import learn
    SP, xf, sx, seeds = _workload(70, 0.4, "int10", 24, 5)
    P = api.make_params()
    mlp = learn.MLP(seed=3)
    try:
        outs = []
        for graphs in (False, True):
            ctx = api.BoundContext(SP, params=P, max_nodes=8)
            ctx.set_precision(slices=4, certify=1)
            api.N.check(api.N.lib().bqb_ctx_set_model(ctx.h, mlp.h, 0.1))
            ctx.set_graphs(graphs)
            outs.append([ctx.bound(xf, sx, LB=0.0, diff=0.0, seeds=seeds, stream=True) for _ in range(2)])
            ctx.close()
        for a, b in zip(outs[0], outs[1]):
            _same(a, b)
    finally:
        mlp.close()


@pytest.mark.parametrize("batch", [1, 8])
def test_solve_with_graphs_same_tree(batch):
    adj = I.gnp_adj(50, 0.5, 7, weights="unit")
    SP = api.maxcut_L(adj)
    P = api.make_params(include_Pent=0, include_Hepta=0)
    st0, sol0 = api.solve(SP, params=P, seed=2020, batch=batch, graphs=False)
    st1, sol1 = api.solve(SP, params=P, seed=2020, batch=batch, graphs=True)
    assert st0.status == 0 and st1.status == 0
    assert st0.best == st1.best and st0.nodes == st1.nodes and st0.root_bound == st1.root_bound
    assert np.array_equal(sol0, sol1)
