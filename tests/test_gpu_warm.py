"""Inherited cuts on the GPU (bqb_ctx_set_warm / bqb_solve_options.warm):
a child node starts its bundle from its parent's final cuts and multipliers.

What pins it:
* the evaluation of the inherited multipliers is fct_eval (src/bundle.c:
  242-275) on the child's subproblem: the device's first center equals the
  oracle's fct_eval on the same cuts and multipliers within 1e-6 relative;
* every bound stays a valid upper bound (weak duality for any gamma >= 0):
  checked against brute force on the child's completions;
* the branch and bound with inherited cuts proves the reference fixtures'
  optima and brute-force optima exactly."""
import os

import numpy as np
import pytest

import oracle_py as O
from fixture_util import brute_maxcut, fixture_value
from parallel_biqbin_maxcut_amd import api, instances as I

pytestmark = pytest.mark.gpu


def reference_params(n0, **kw):
    p = api.make_params(**kw)
    if p.adjust_TriIneq:
        p.TriIneq = 10 * n0                  # src/process_input.c:135-136
    return p


def child_cuts(cuts, B, ic):
    """The parent's cuts (subproblem of the root: vertices 0..B-1, then the
    last vertex B) with multiplier >= 1e-5, renumbered for the child that
    fixes variable ic; cuts containing ic are left out."""
    pos = np.arange(B + 1)
    pos[ic] = -1
    pos[ic + 1:] -= 1
    tris, pents, hepts, gam = [], [], [], []
    nt, npn = len(cuts["tri"]), len(cuts["pent"])
    for c, r in enumerate(cuts["tri"]):
        y = cuts["gamma"][c]
        v = pos[r[:3]]
        if y >= 1e-5 and np.all(v >= 0):
            tris.append((int(v[0]), int(v[1]), int(v[2]), int(r[3])))
            gam.append(y)
    for kind, arr, off, k in (("pent", cuts["pent"], nt, 5), ("hept", cuts["hept"], nt + npn, 7)):
        for c, r in enumerate(arr):
            y = cuts["gamma"][off + c]
            v = pos[r[1:1 + k]]
            if y >= 1e-5 and np.all(v >= 0):
                (pents if kind == "pent" else hepts).append((int(r[0]), [int(a) for a in v]))
    # multipliers in [tri | pent | hept] order
    for kind, arr, off, k in (("pent", cuts["pent"], nt, 5), ("hept", cuts["hept"], nt + npn, 7)):
        for c, r in enumerate(arr):
            y = cuts["gamma"][off + c]
            if y >= 1e-5 and np.all(pos[r[1:1 + k]] >= 0):
                gam.append(y)
    return tris, pents, hepts, np.array(gam)


def root_with_cuts(SP, P, seed=5):
    """The root bound of a warm context; the parameters stop it after few
    outer rounds (a give-up), so that it ends with cuts instead of a prune."""
    B = SP.shape[0] - 1
    ctx = api.BoundContext(SP, params=P, max_nodes=2)
    ctx.set_warm(True)
    z = np.zeros(B, np.int32)
    ctx.push(z, z, LB=0.0, diff=0.0, is_root=1, seed=seed)
    while ctx.busy():
        ctx.advance()
    root = ctx.pop()
    assert root["status"] == 0 and root["stats"][0] == 0 and root["stats"][1] == 1
    return ctx, root


def test_inherited_center_is_fct_eval_of_the_inherited_multipliers():
    n0 = 40
    SP = O.build_L(I.gnp_adj(n0, 0.5, 11, weights="pm1"))
    P = reference_params(n0, max_outer_iter=3, extra_iter=0)
    P.include_Pent = P.include_Hepta = 1
    P.triag_iter = 1                         # k-gonal separation from round 2 on
    ctx, root = root_with_cuts(SP, P)
    B = n0 - 1
    cuts = root["cuts"]
    assert len(cuts["tri"]) > 0 and np.all(cuts["gamma"] >= -1e-12)
    ic = int(np.argmin(np.abs(0.5 - root["fracsol"])))
    checked = 0
    for v in (0, 1):
        xf = np.zeros(B, np.int32)
        sx = np.zeros(B, np.int32)
        xf[ic], sx[ic] = 1, v
        tris, pents, hepts, gam = child_cuts(cuts, B, ic)
        assert len(tris) > 0
        slot = ctx.push(xf, sx, LB=0.0, diff=1e9, is_root=0, seed=9 + v, tag=v)
        ctx.push_cuts(slot, tris, [[t] + p for t, p in pents], [[t] + p for t, p in hepts], gam)
        ctx.run(api.PH["BASIC_TRAIN"])
        st = ctx.state(slot)
        assert st.status == 0 and st.active == 1 and st.warm == 0 and st.bundle == 2
        # the inherited cuts lead the device's lists, unchanged
        m = len(tris)
        dev_tri = ctx.get("tri", slot, np.int32, 4 * m).reshape(m, 4)
        assert np.array_equal(dev_tri, np.array(tris, np.int32))
        assert st.NPentIneq == len(pents) and st.NHeptaIneq == len(hepts) and st.NIneq >= m
        PL, fv = O.create_subproblem(SP, xf, sx)
        fo, _, _, ost = O.fct_eval(PL, O.cut_arrays(tris, pents, hepts), gam)
        assert ost.status == 0
        phi0 = st.basic_bound - fv
        if fo < phi0:
            assert abs(st.f - fo) <= 1e-6 * abs(fo), (st.f, fo)
            gdev = ctx.get("gamma", slot, np.float64, len(gam))
            assert np.array_equal(gdev[:m], gam[:m])
            checked += 1
        else:
            assert st.f == phi0
        while ctx.busy():
            ctx.advance()
        ch = ctx.pop()
        assert ch["status"] == 0
    assert checked >= 1
    ctx.close()


@pytest.mark.parametrize("seed", [3, 4])
def test_inherited_child_bounds_are_valid(seed):
    """Both children of the root of an 18-vertex graph, with and without the
    root's cuts: every bound is at least the best completion (brute force)."""
    n0 = 18
    adj = I.gnp_adj(n0, 0.5, seed, weights="int10")
    SP = api.maxcut_L(adj)
    P = reference_params(n0, max_outer_iter=1, extra_iter=0)
    ctx, root = root_with_cuts(SP, P, seed)
    B = n0 - 1
    cuts = root["cuts"]
    ic = int(np.argmin(np.abs(0.5 - root["fracsol"])))
    bits = np.arange(1 << B)[:, None]
    X = ((bits >> np.arange(B)) & 1).astype(np.int64)
    vals = np.einsum("ki,ij,kj->k", X, SP[:B, :B], X)
    for v in (0, 1):
        xf = np.zeros(B, np.int32)
        sx = np.zeros(B, np.int32)
        xf[ic], sx[ic] = 1, v
        best = vals[X[:, ic] == v].max()
        tris, pents, hepts, gam = child_cuts(cuts, B, ic)
        slot = ctx.push(xf, sx, LB=0.0, diff=1e9, is_root=0, seed=21 + v, tag=v)
        if len(gam):
            ctx.push_cuts(slot, tris, [[t] + p for t, p in pents], [[t] + p for t, p in hepts], gam)
        while ctx.busy():
            ctx.advance()
        ch = ctx.pop()
        assert ch["status"] == 0
        assert ch["bound"] >= best - 1e-9, (v, ch["bound"], best)
    ctx.close()


def test_push_cuts_rejects_invalid_records():
    n0 = 20
    SP = O.build_L(I.gnp_adj(n0, 0.5, 2, weights="pm1"))
    ctx = api.BoundContext(SP, params=reference_params(n0), max_nodes=2)
    B = n0 - 1
    z = np.zeros(B, np.int32)
    with pytest.raises(RuntimeError):            # warm is off
        slot = ctx.push(z, z, tag=0)
        ctx.push_cuts(slot, [(3, 2, 1, 1)], gamma=[0.5])
    ctx.close()
    ctx = api.BoundContext(SP, params=reference_params(n0), max_nodes=2)
    ctx.set_warm(True)
    slot = ctx.push(z, z, tag=0)                  # 19 free variables + the last vertex: n = 20
    for bad, g in (([(20, 2, 1, 1)], [0.5]),      # vertex outside the subproblem
                   ([(1, 2, 3, 1)], [0.5]),       # not i > j > k
                   ([(3, 2, 1, 5)], [0.5]),       # type
                   ([(3, 2, 1, 1)], [-1.0]),      # negative multiplier
                   ([(3, 2, 1, 1)], [np.nan])):
        with pytest.raises(RuntimeError):
            ctx.push_cuts(slot, bad, gamma=g)
    with pytest.raises(RuntimeError):            # repeated pentagonal vertex
        ctx.push_cuts(slot, pents=[[1, 0, 1, 2, 3, 3]], gamma=[0.5])
    ctx.push_cuts(slot, [(19, 2, 1, 4)], gamma=[0.5])
    while ctx.busy():
        ctx.advance()
    assert ctx.pop()["status"] == 0
    ctx.close()


@pytest.mark.parametrize("name,batch", [("g05_60.0", 1), ("g05_80.0", 1), ("g05_80.0", 16), ("g05_100.4", 16)])
def test_warm_bab_proves_fixture_optima(golden, name, batch):
    path = os.path.join(golden, "rudy", name)
    SP = api.maxcut_L(api.read_rudy(path))
    st, sol = api.solve(SP, reference_params(SP.shape[0]), seed=2020, batch=batch, warm=True)
    assert st.status == 0 and not st.stopped
    assert st.best == fixture_value(path + "-expected_output")
    assert api.evaluate_solution(SP, sol) == st.best


@pytest.mark.parametrize("seed,batch", [(1, 1), (2, 8), (5, 4)])
def test_warm_bab_brute_force(seed, batch):
    adj = I.gnp_adj(17, 0.5, seed, weights="int10")
    SP = api.maxcut_L(adj)
    st, sol = api.solve(SP, reference_params(17), seed=seed, batch=batch, warm=True)
    assert st.status == 0
    assert st.best == brute_maxcut(adj)
    assert api.evaluate_solution(SP, sol) == st.best


def test_warm_with_certified_mixed_bounds_and_graphs():
    """Inherited cuts together with bf16 matrix-core products, certified
    bounds and launch trains as HIP graphs: the same proven optimum."""
    SP = api.maxcut_L(I.gnp_adj(60, 0.5, 19, weights="unit"))
    P = api.make_params(include_Pent=0, include_Hepta=0)
    P.TriIneq = 600
    plain, _ = api.solve(SP, P, seed=2020, batch=1)
    for kw in (dict(slices=4, certify=1), dict(graphs=True)):
        st, sol = api.solve(SP, P, seed=2020, batch=16, warm=True, **kw)
        assert st.status == 0 and st.best == plain.best
        assert api.evaluate_solution(SP, sol) == st.best
