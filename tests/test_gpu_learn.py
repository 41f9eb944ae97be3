"""The learned cut model on the GPU: the library's own bf16 forward and
backward kernels against the CPU FP64 oracle (oracle/orc_mlp.py), the
device feature extractor against its specification (learn.cut_features),
training, solver-generated data, and the model inside the node bound
(bounds stay valid: any multiplier >= 0 gives a valid bound).

Tolerances: features within 1e-5 relative (FP32 on the device); gradients
within 2 % relative norm error per parameter block and cosine >= 0.999,
loss within 1e-4 relative (FP32 accumulation against FP64 with the same
bf16 roundings)."""
import numpy as np
import pytest

import oracle_py as O
import orc_mlp as M
from fixture_util import brute_maxcut
from parallel_biqbin_maxcut_amd import api, instances as I, learn

pytestmark = pytest.mark.gpu


def _rand_node(n, seed):
    rng = np.random.default_rng(seed)
    V = rng.normal(size=(n, 4))
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    X = V @ V.T
    A = np.triu(rng.integers(-3, 4, size=(n, n)).astype(float), 1)
    L = A + A.T
    np.fill_diagonal(L, rng.integers(0, 5, size=n))
    return X, L / 4.0


def _rand_cuts(n, seed):
    rng = np.random.default_rng(seed)
    cuts = []
    for t in range(1, 5):
        for _ in range(3):
            i, j, k = sorted(rng.choice(n, 3, replace=False))[::-1]
            cuts.append((3, t, [int(i), int(j), int(k)]))
    for t in range(1, 4):
        cuts.append((5, t, [int(v) for v in rng.choice(n, 5, replace=False)]))
    for t in range(1, 5):
        cuts.append((7, t, [int(v) for v in rng.choice(n, 7, replace=False)]))
    return cuts


@pytest.mark.parametrize("n", [12, 40])
def test_device_features_match_spec(n):
    X, L = _rand_node(n, n)
    cuts = _rand_cuts(n, n + 1)
    dev, order = learn.device_features(X, L, cuts, 3.0, 0.125)
    ref = learn.cut_features(X, L, order, 3.0, 0.125)
    assert np.allclose(dev, ref, rtol=1e-5, atol=1e-6), np.max(np.abs(dev - ref))


def _blocks():
    o = 0
    for name, shp in M.SIZES:
        k = int(np.prod(shp))
        yield name, o, k
        o += k


def test_gpu_gradient_matches_cpu_oracle():
    rng = np.random.default_rng(1)
    mlp = learn.MLP(seed=5)
    try:
        p = mlp.params().astype(np.float64)
        x = np.zeros((300, 32), dtype=np.float32)
        x[:, :16] = rng.normal(size=(300, 16))
        t = np.stack([np.abs(rng.normal(size=300)) * 2, (rng.random(300) < 0.4)], axis=1).astype(np.float32)
        g, loss = mlp.gradient(x, t)
        lref, gref = M.loss_and_grad(p, x.astype(np.float64), t.astype(np.float64), bf=True)
        print(f"loss gpu {loss:.8f} cpu {lref:.8f}", flush=True)
        assert abs(loss - lref) <= 1e-4 * abs(lref)
        for name, o, k in _blocks():
            a, b = g[o:o + k].astype(np.float64), gref[o:o + k]
            nb = np.linalg.norm(b)
            if nb == 0:
                assert np.all(a == 0)
                continue
            rel = np.linalg.norm(a - b) / nb
            cos = a @ b / (np.linalg.norm(a) * nb)
            print(f"{name}: rel {rel:.2e} cos {cos:.6f}", flush=True)
            assert rel <= 2e-2 and cos >= 0.999, name
    finally:
        mlp.close()


def test_training_reduces_loss():
    rng = np.random.default_rng(2)
    x = np.zeros((20000, 32), dtype=np.float32)
    x[:, :16] = rng.normal(size=(20000, 16))
    t = np.stack([np.log1p(np.abs(x[:, 0] * x[:, 1])), (x[:, 2] + 0.5 * x[:, 3] > 0)], axis=1).astype(np.float32)
    mlp = learn.MLP(seed=1)
    try:
        losses = mlp.train(x, t, epochs=8, batch=1024, lr=2e-3, seed=3)
        print("losses", losses, flush=True)
        assert losses[-1] < 0.5 * losses[0]
        pred = mlp.forward(x[:2000])
        assert np.mean((pred[:, 1] > 0) == (t[:2000, 1] > 0.5)) > 0.9
    finally:
        mlp.close()


def test_solver_generated_records():
    P = api.make_params()
    P.TriIneq = 400
    graphs = [I.gnp_adj(40, 0.5, s, weights="zpm1") for s in range(3)]
    feat, target = learn.generate(graphs, P, seed=1, workers=3, cap=20000)
    print(f"records {len(feat)}, active {target[:, 1].mean():.3f}", flush=True)
    assert len(feat) > 100
    assert np.all(np.isfinite(feat)) and np.all(target[:, 0] >= 0)
    assert 0.0 < target[:, 1].mean() < 1.0
    assert np.all(feat[:, 15] == 1.0) and np.all(feat[:, 16:] == 0.0)
    # every recorded triangle was violated by more than the threshold when it
    # was appended (B(X) - 1 > violated_TriIneq)
    tri = feat[:, 2] == 1.0
    assert tri.any() and np.all(feat[tri, 1] > P.violated_TriIneq - 1e-6)


@pytest.mark.parametrize("drop", [0.0, 0.3])
def test_model_in_bound_gives_valid_bounds(drop):
    """A model (random weights: arbitrary multipliers >= 0) inside the bound:
    the bounds are still valid and the branch and bound is unaffected."""
    mlp = learn.MLP(seed=11)
    try:
        for s in range(4):
            adj = I.gnp_adj(20, 0.5, s, weights="int10")
            SP = api.maxcut_L(adj)
            P = api.make_params()
            P.TriIneq = 200
            ctx = api.BoundContext(SP, params=P, max_nodes=1)
            api.N.check(api.N.lib().bqb_ctx_set_model(ctx.h, mlp.h, drop))
            z = np.zeros((1, 19), np.int32)
            r = ctx.bound(z, z, LB=0.0, is_root=1, seeds=np.uint64(s))
            ctx.close()
            assert r["status"][0] == 0
            assert r["bound"][0] >= brute_maxcut(adj) - 1e-9
    finally:
        mlp.close()


def _round_to_rebuild(ctx, nb):
    """One outer round of a one-slot context, phase by phase as the launch
    train runs it, up to (not including) the rebuild."""
    PH = api.PH
    ctx.run(PH["OUTER_BEGIN"])
    for b in range(nb):
        ctx.run(PH["BUNDLE_NEXT"], b)
        ctx.run(PH["QP"], 0)
        ctx.run(PH["OPBT"], 1)
        ctx.run(PH["IPM_L0"])
        ctx.run(PH["STEP"])
    for ph in ("OUTER_A", "HEUR", "OUTER_B", "TRISEP", "SEP_PENT", "PENT", "SEP_HEPT", "HEPT"):
        ctx.run(PH[ph])
    ctx.run(113)                       # the model on the round's new cuts
    for ph, arg in (("OUTER_C", 0), ("CSR", 0), ("OPBT", 0), ("IPM_L0", 0)):
        ctx.run(PH[ph], arg)


@pytest.mark.parametrize("kind", ["trained", "huge"])
def test_learned_warm_start_kept_only_where_it_lowers_f(kind):
    """The rebuild after a separation evaluates the model's multipliers of
    the new cuts; the center moves to them only where that lowers f.  A
    model biased to predict multipliers near 20 for every cut raises f:
    the center and f stay and the new cuts start at 0, as in the reference
    (bounding.c:278 extends the bundle only).  The trained model never
    raises f at a rebuild either."""
    import os
    path = os.path.join(os.path.dirname(learn.__file__), "models", "cut_model.npz")
    prm = np.array(np.load(path, allow_pickle=False)["params"], dtype=np.float32)
    mlp = learn.MLP(params=prm)
    n0 = 40
    SP = api.maxcut_L(I.gnp_adj(n0, 0.5, 3, weights="unit"))
    P = api.make_params(include_Pent=0, include_Hepta=0)
    P.TriIneq = 10 * n0
    ctx = api.BoundContext(SP, params=P, max_nodes=1)
    try:
        api.N.check(api.N.lib().bqb_ctx_set_model(ctx.h, mlp.h, 0.0))
        z = np.zeros(n0 - 1, np.int32)
        slot = ctx.push(z, z, LB=0.0, diff=0.0, is_root=1, seed=5, tag=0)
        ctx.run(api.PH["BASIC_TRAIN"])
        if kind == "huge":
            # from the first round end on (the first separation's cuts stay
            # as the trained model set them, so that later rounds add cuts)
            b4 = learn.NPARAM - 16         # the output layer's 16 biases (rows 0 and 1 used)
            prm[b4] = prm[b4 + 1] = 100.0  # magnitude at its cap, P(active) = 1
            mlp.set_params(prm)
        for r in range(2):
            st = ctx.state(slot)
            assert st.status == 0 and st.active == 1 and st.rphase == 0
            _round_to_rebuild(ctx, P.init_bundle_iter if r == 0 else P.init_bundle_iter + 1)
            before = ctx.state(slot)
            assert before.sep_tri == 1 and before.nsurv_tri < before.NIneq     # new cuts exist
            m = before.NIneq
            g0 = ctx.get("gamma", slot, np.float64, m)
            ctx.run(api.PH["REBUILD"])
            after = ctx.state(slot)
            g1 = ctx.get("gamma", slot, np.float64, m)
            assert after.status == 0 and after.f <= before.f
            assert np.array_equal(g1[:before.nsurv_tri], g0[:before.nsurv_tri])   # survivors untouched
            if kind == "huge":
                assert np.all(g0[before.nsurv_tri:] > 1.0)                         # predictions made
                assert after.f == before.f and np.all(g1[before.nsurv_tri:] == 0.0)
            elif after.f < before.f:
                assert np.array_equal(g1, g0)                                      # the warm start kept
    finally:
        ctx.close()
        mlp.close()


def test_generate_and_train_from_solver_data():
    """The whole loop on a small scale: solver-generated pairs from 40
    G(80, p) graphs, trained on the GPU; the held-out loss falls below the
    loss of the untrained model."""
    P = api.make_params()
    P.TriIneq = 800
    feat, target = learn.generate(learn.training_graphs(40, 99), P, seed=5, workers=8)
    n = len(feat)
    tr, va = np.arange(n) < int(0.8 * n), np.arange(n) >= int(0.8 * n)
    mlp = learn.MLP(seed=2)
    try:
        _, before = mlp.gradient(feat[va], target[va])
        mlp.train(feat[tr], target[tr], epochs=3, batch=1024, lr=1e-3, seed=1)
        _, after = mlp.gradient(feat[va], target[va])
        print(f"{n} rows; held-out loss {before:.4f} -> {after:.4f}", flush=True)
        assert after < 0.8 * before
    finally:
        mlp.close()


def test_trained_model_gain_on_heldout_graphs():
    """cfg5: the model trained on 2000 G(80, p ~ U[0.1, 0.9]) graphs with
    weights in {-1, 0, +1} (parallel_biqbin_maxcut_amd/models/cut_model.npz,
    written by train_cut_model.py) against the plain cold-start bundle on
    100 held-out G(120, 0.5) +-1 graphs, root bound, 6 outer rounds.  The
    plain run is the judge: every learned bound is valid (at least the best
    cut found); the gain in bound gap and IPM iterations is reported and
    must not be negative on average."""
    import json
    import os
    path = os.path.join(os.path.dirname(learn.__file__), "models", "cut_model.npz")
    d = np.load(path, allow_pickle=False)
    mlp = learn.MLP(params=d["params"])
    try:
        P = api.make_params(max_outer_iter=6, extra_iter=0)
        P.TriIneq = 1200
        per, s = learn.evaluate_gain(mlp, learn.eval_graphs(100), P, drop=0.0, workers=16)
        print(json.dumps(s), flush=True)
        assert s["status_ok"]
        assert np.all(per["bound_learned"] >= per["lb"] - 1e-9)
        assert s["mean_gap_learned"] <= s["mean_gap_plain"]
        # the warm start acts in the early rounds
        g = s["gap_by_round_plain_learned"]
        assert g[1][1] < g[1][0]
    finally:
        mlp.close()
