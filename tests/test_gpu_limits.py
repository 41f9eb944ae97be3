"""Capacity limits and their handling on the GPU path.

* A node whose bundle reaches the context's capacity ends as a give-up with
  the bound of its stability center (src/bounding.c would abort the run at
  MaxBundle, :312-315); the node queue branches it, so the optimum is still
  proven, and the event is counted.
* A separation that produces more cuts than the context holds is flagged.
* Round 1 of the outer loop runs init_bundle_iter bundle steps even when it
  exceeds max_bundle_iter (src/bounding.c:50, 323-325)."""
import os

import numpy as np
import pytest

import oracle_py as O
from fixture_util import brute_maxcut, fixture_value
from parallel_biqbin_maxcut_amd import api, instances as I

pytestmark = pytest.mark.gpu


def _params(n0, **kw):
    p = api.make_params(**kw)
    if p.adjust_TriIneq:
        p.TriIneq = 10 * n0
    return p


@pytest.mark.parametrize("seed", [4, 5])
def test_bundle_full_gives_up_and_optimum_is_still_proven(seed):
    adj = I.gnp_adj(22, 0.5, seed, weights="int10")
    SP = api.maxcut_L(adj)
    st, sol = api.solve(SP, _params(22), seed=seed, batch=16, max_bundle=3)
    print(f"seed {seed}: nodes {st.nodes}, bundle-full give-ups {st.bundle_full}", flush=True)
    assert st.status == 0 and not st.stopped
    assert st.bundle_full > 0
    assert st.best == brute_maxcut(adj)
    assert api.evaluate_solution(SP, sol) == st.best


def test_bundle_full_fixture_g05_60(golden):
    path = os.path.join(golden, "rudy", "g05_60.0")
    SP = api.maxcut_L(api.read_rudy(path))
    st, sol = api.solve(SP, _params(SP.shape[0]), seed=2020, batch=64, max_bundle=6)
    print(f"g05_60 max_bundle 6: nodes {st.nodes}, bundle-full give-ups {st.bundle_full}", flush=True)
    assert st.status == 0 and st.bundle_full > 0
    assert st.best == fixture_value(path + "-expected_output")


def test_bundle_full_node_bound_is_valid():
    """The given-up node's bound is the bound of its center at that moment:
    between the node's own cut and the basic SDP bound."""
    n0 = 60
    SP = O.build_L(I.gnp_adj(n0, 0.5, 1, weights="unit"))
    P = _params(n0, include_Pent=0, include_Hepta=0)
    xf = np.zeros((1, n0 - 1), np.int32)
    ctx = api.BoundContext(SP, params=P, max_nodes=1, max_bundle=3)
    r = ctx.bound(xf, xf, LB=0.0, diff=0.0, is_root=1, seeds=5)
    ctx.close()
    o, _, _ = O.sdpbound(SP, xf[0], xf[0], 0.0, 0.0, 1, 5, O.default_params(**{
        f: getattr(P, f) for f, _ in api.BiqBinParameters._fields_}))
    assert r["status"][0] == 0
    assert r["stats"][0][8] & api.FLAG_BUNDLE_FULL
    assert r["stats"][0][1] == 1                       # give-up
    _, _, phi0, _, _ = O.ipm(O.create_subproblem(SP, xf[0], xf[0])[0])
    assert r["LB"][0] <= r["bound"][0] <= phi0 * (1 + 1e-9)
    assert o.bound <= phi0 * (1 + 1e-9)


def test_triangle_truncation_is_flagged():
    n0 = 80
    SP = O.build_L(I.gnp_adj(n0, 0.5, 2, weights="pm1"))
    P = _params(n0, include_Pent=0, include_Hepta=0, max_outer_iter=3, extra_iter=0)
    xf = np.zeros((1, n0 - 1), np.int32)
    ctx = api.BoundContext(SP, params=P, max_nodes=1, max_tri=64)
    r = ctx.bound(xf, xf, LB=0.0, diff=0.0, is_root=1, seeds=5)
    ctx.close()
    assert r["status"][0] == 0
    assert r["stats"][0][8] & api.FLAG_TRI_TRUNC
    assert r["stats"][0][5] <= 64


def test_round_one_runs_init_bundle_iter_uncapped():
    """init_bundle_iter = 6 > max_bundle_iter = 2: the reference's round 1
    runs 6 bundle steps, later rounds 2; same decisions and bound as the
    oracle (which follows bounding.c line by line)."""
    n0 = 60
    SP = O.build_L(I.gnp_adj(n0, 0.5, 1, weights="unit"))
    P = _params(n0, include_Pent=0, include_Hepta=0, max_outer_iter=3, extra_iter=0,
                init_bundle_iter=6, max_bundle_iter=2)
    xf = np.zeros((1, n0 - 1), np.int32)
    ctx = api.BoundContext(SP, params=P, max_nodes=1)
    r = ctx.bound(xf, xf, LB=0.0, diff=0.0, is_root=1, seeds=5)
    ctx.close()
    o, _, _ = O.sdpbound(SP, xf[0], xf[0], 0.0, 0.0, 1, 5, O.default_params(**{
        f: getattr(P, f) for f, _ in api.BiqBinParameters._fields_}))
    s = r["stats"][0]
    assert r["status"][0] == 0 and o.status == 0
    assert (s[0], s[1], s[2], s[3]) == (o.prune, o.giveup, o.outer_iters, o.ipm_calls)
    assert abs(r["bound"][0] - o.bound) <= 1e-6 * abs(o.bound), (r["bound"][0], o.bound)


def test_reference_cut_capacities():
    """Contexts holding the reference's cut capacities (MaxTriIneqAdded
    50 000, pentagonal and heptagonal 50 000, src/biqbin.h:37-39) bound nodes
    with results bitwise equal to the default capacities' when nothing
    overflows (capacities change strides, not arithmetic)."""
    n0 = 60
    SP = O.build_L(I.gnp_adj(n0, 0.5, 8, weights="pm1"))
    P = _params(n0, max_outer_iter=4, extra_iter=0)
    xf, sx = I.random_fixings(n0, 8, 3, 9)
    seeds = np.arange(1, 4, dtype=np.uint64)
    out = []
    for caps in ((0, 0, 0), (50000, 50000, 50000)):
        ctx = api.BoundContext(SP, params=P, max_nodes=3, max_tri=caps[0], max_pent=caps[1], max_hept=caps[2])
        r = ctx.bound(xf, sx, LB=0.0, diff=1e9, is_root=0, seeds=seeds)
        ctx.close()
        assert np.all(r["status"] == 0) and not np.any(r["stats"][:, 8] & (api.FLAG_TRI_TRUNC | api.FLAG_PENT_TRUNC
                                                                        | api.FLAG_HEPT_TRUNC))
        out.append(r)
    assert np.array_equal(out[0]["bound"], out[1]["bound"])
    assert np.array_equal(out[0]["stats"], out[1]["stats"])
