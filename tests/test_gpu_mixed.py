"""The mixed-precision form on the GPU: the IPM products on the bf16 matrix
cores (sliced operands, exact FP32 accumulation, FP64 combination), FP64
inversions and Cholesky line searches, and certified bounds computed with
directed rounding.

What is compared, and at what tolerance:
* the sliced product itself against an FP64 product, within its a priori
  error bound (4s + 1) n 2^-8s max|A(i,:)| max|B(:,j)| (plus FP64 rounding);
* the mixed IPM's phi (4 slices, the setting the solver uses) against the
  CPU oracle's FP64 phi on identical inputs, within 1e-6 relative (the
  per-node bound criterion), per call (3 slices leave up to 2e-6 and can
  stall a line search, which the FP64 re-solve then covers);
* certified bounds: never below the oracle's FP64 value minus 1e-6
  relative, at most 1e-7 relative above the uncertified value, never below
  the brute-force maximum cut (n <= 22), and, for n <= 10, exactly valid:
  Diag(y) - L + ((phi_cert - e'y)/n) I is positive semidefinite in rational
  arithmetic;
* whole nodes of G(250, 0.1) with weights +-1..+-10 (three outer rounds,
  triangle cuts): each node's certified mixed bound against the FP64
  oracle's, error stated, within 1e-4 relative (the trajectory's own
  sensitivity, see test_gpu_bound.py) and never below the node's cut;
* branch and bound with certified mixed bounds proves the reference's
  optima exactly."""
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle_py as O
from fixture_util import brute_maxcut, fixture_value
from parallel_biqbin_maxcut_amd import api, instances as I

pytestmark = pytest.mark.gpu


def _subs(n0, p, weights, seed, count, nfix):
    L = O.build_L(I.gnp_adj(n0, p, seed, weights=weights))
    xf, sx = I.random_fixings(n0, nfix, count, seed + 1)
    return [O.create_subproblem(L, xf[b], sx[b])[0] for b in range(count)]


@pytest.mark.parametrize("n,s", [(16, 3), (40, 2), (100, 3), (151, 3), (250, 3), (256, 4)])
def test_sliced_product_within_its_error_bound(n, s):
    rng = np.random.default_rng(n + s)
    A = rng.normal(size=(n, n)) * np.exp(rng.normal(size=(n, 1)) * 3)
    B = rng.normal(size=(n, n)) * np.exp(rng.normal(size=(1, n)) * 3)
    A[rng.random((n, n)) < 0.2] = 0.0
    C = api.mx_product(A, B, slices=s)
    ref = A @ B
    scale = np.abs(A).max(axis=1)[:, None] * np.abs(B).max(axis=0)[None, :]
    bound = ((4 * s + 1) * n * 2.0 ** (-8 * s) + 4 * n * 2.0 ** -53) * scale
    err = np.abs(C - ref)
    print(f"n {n} s {s}: max err/bound {np.max(err / bound):.3e}", flush=True)
    assert np.all(err <= bound + 1e-300)


@pytest.mark.parametrize("slices", [4])
def test_mixed_ipm_phi_matches_fp64_oracle(slices):
    Ls = (_subs(100, 0.3, "pm1", 3, 8, 40) + _subs(151, 0.5, "int10", 5, 3, 40)
          + _subs(250, 0.1, "int10", 31, 3, 40) + _subs(30, 0.5, "unit", 2, 3, 5))
    r = api.ipm_batch_mixed(Ls, slices=slices)
    worst = 0.0
    for b, L in enumerate(Ls):
        _, _, phi, st, _ = O.ipm(L)
        assert r["status"][b] == 0 and st.status == 0
        rel = abs(r["phi"][b] - phi) / abs(phi)
        worst = max(worst, rel)
        print(f"n {L.shape[0]}: phi rel err {rel:.2e}, iters {r['iters'][b]} vs {st.iters}", flush=True)
        assert rel <= 1e-6, (b, L.shape[0], r["phi"][b], phi)
        cert = r["phi_cert"][b]
        assert cert >= phi * (1 - 1e-6)
        assert cert >= r["phi"][b]
        assert cert - r["phi"][b] <= 1e-7 * abs(r["phi"][b]), (cert, r["phi"][b])
    print(f"slices {slices}: worst phi error {worst:.2e}", flush=True)


@pytest.mark.parametrize("slices", [0, 4])
def test_certified_phi_above_brute_force(slices):
    for seed in range(6):
        adj = I.gnp_adj(18, 0.5, seed, weights="int10")
        L, _ = O.create_subproblem(O.build_L(adj), np.zeros(17, np.int32), np.zeros(17, np.int32))
        r = api.ipm_batch_mixed([L], slices=slices)
        assert r["status"][0] == 0
        assert r["phi_cert"][0] >= brute_maxcut(adj)


def _psd_exact(M):
    """Exact positive semidefiniteness of a symmetric rational matrix by
    symmetric Gaussian elimination (LDL')."""
    A = [row[:] for row in M]
    n = len(A)
    for k in range(n):
        piv = A[k][k]
        if piv < 0:
            return False
        if piv == 0:
            if any(A[k][j] != 0 for j in range(k + 1, n)):
                return False
            continue
        for i in range(k + 1, n):
            f = A[i][k] / piv
            if f != 0:
                for j in range(k + 1, n):
                    A[i][j] -= f * A[k][j]
    return True


@pytest.mark.parametrize("slices", [0, 4])
def test_certificate_is_exactly_valid_small(slices):
    """Diag(y) - L + ((phi_cert - e'y)/n) I is psd in exact arithmetic, with
    y and phi_cert the doubles the GPU returned and L exact."""
    Ls = _subs(10, 0.6, "int10", 7, 4, 2)
    r = api.ipm_batch_mixed(Ls, slices=slices)
    for b, L in enumerate(Ls):
        n = L.shape[0]
        assert r["status"][b] == 0
        y = [Fraction(v) for v in r["y"][b]]
        sh = (Fraction(r["phi_cert"][b]) - sum(y)) / n
        Z = [[(y[i] if i == j else 0) - Fraction(L[i, j]) + (sh if i == j else 0) for j in range(n)]
             for i in range(n)]
        assert _psd_exact(Z), b


def _bt_pattern(n, cut):
    """B'(e_c) of one triangle cut, exactly: the oracle's op_Bt of a zero
    matrix with multiplier 1 gives -B'(1), entries 0 or +-1/2."""
    X = np.zeros((n, n), order="F")
    X = O.op_Bt(X, O.cut_arrays([cut]), np.ones(1))
    return [[-Fraction(X[i, j]) for j in range(n)] for i in range(n)]


@pytest.mark.parametrize("slices", [0, 4])
def test_round_end_certificate_is_exactly_valid(slices):
    """The certificate computed at the end of each outer round for the
    bundle center (y of the center's evaluation, its multipliers and cuts)
    is exactly valid: Diag(y) - L + B'gamma + ((f_cert - sum(gamma) - e'y)/n) I
    is psd in rational arithmetic.  The node runs phase by phase so the
    check sees the cut set before the round's separation changes it."""
    n0 = 14
    SP = O.build_L(I.gnp_adj(n0, 0.6, 5, weights="int10"))
    P = api.make_params(include_Pent=0, include_Hepta=0)
    P.TriIneq = 140
    ctx = api.BoundContext(SP, params=P, max_nodes=1)
    ctx.set_precision(slices=slices, certify=1)
    z = np.zeros(n0 - 1, np.int32)
    ctx.push(z, z, LB=0.0, diff=1e9, is_root=0, seed=3)
    ctx.run(api.PH["BASIC_TRAIN"])
    PL, fv = O.create_subproblem(SP, z, z)
    n = PL.shape[0]
    checked = 0
    for rnd in range(3):
        st = ctx.state(0)
        if not st.active:
            break
        ctx.run(api.PH["OUTER_BEGIN"])
        for b in range(ctx.state(0).bdl_left):
            ctx.run(api.PH["BUNDLE_NEXT"], b)
            ctx.run(api.PH["QP"], 0)
            ctx.run(api.PH["OPBT"], 1)
            ctx.run(api.PH["IPM_L0"])
            ctx.run(api.PH["STEP"])
        stale = ctx.state(0).cert_stale
        ctx.run(api.PH["CERT_MASK"])
        ctx.run(111)
        st = ctx.state(0)
        assert st.cert_stale == 0 and st.status == 0
        if stale:
            m = st.NIneq
            y = [Fraction(v) for v in ctx.get("ycen", 0, np.float64, n)]
            # the certificate is that of max(gamma, 0) (rounding-level
            # negatives of the bundle's multipliers are dropped)
            gam = np.maximum(ctx.get("gamma", 0, np.float64, m), 0.0)
            tri = ctx.get("tri", 0, np.int32, 4 * m).reshape(m, 4)
            Z = [[(y[i] if i == j else 0) - Fraction(PL[i, j]) for j in range(n)] for i in range(n)]
            for c in range(m):
                if gam[c] == 0:
                    continue
                pat = _bt_pattern(n, tuple(int(v) for v in tri[c]))
                g = Fraction(gam[c])
                for i in range(n):
                    for j in range(n):
                        if pat[i][j]:
                            Z[i][j] += g * pat[i][j]
            sh = (Fraction(st.f_cert) - sum(Fraction(v) for v in gam) - sum(y)) / n
            for i in range(n):
                Z[i][i] += sh
            assert _psd_exact(Z), (rnd, st.f_cert)
            assert st.f_cert >= st.f                       # certified value at least f
            checked += 1
        # rest of the round: heuristic, tests, separation, rebuild
        for ph in ("OUTER_A", "HEUR", "OUTER_B", "TRISEP", "SEP_PENT", "PENT", "SEP_HEPT", "HEPT", "OUTER_C",
                   "CSR"):
            ctx.run(api.PH[ph])
        ctx.run(api.PH["OPBT"], 0)
        ctx.run(api.PH["IPM_L0"])
        ctx.run(api.PH["REBUILD"])
    ctx.close()
    assert checked >= 1


def _bound_both(SP, P, xf, sx, seeds, slices):
    n0 = SP.shape[0]
    ctx = api.BoundContext(SP, params=P, max_nodes=xf.shape[0])
    ctx.set_precision(slices=slices, certify=1)
    r = ctx.bound(xf, sx, LB=0.0, diff=0.0, is_root=0, seeds=seeds)
    ctx.close()
    d = {f: getattr(P, f) for f, _ in api.BiqBinParameters._fields_}
    if d["adjust_TriIneq"]:
        d["TriIneq"] = 10 * n0
    outs = [O.sdpbound(SP, xf[b], sx[b], 0.0, 0.0, 0, int(seeds[b]), O.default_params(**d))[0]
            for b in range(xf.shape[0])]
    return r, outs


@pytest.mark.parametrize("slices", [4])
def test_cfg3_nodes_mixed_certified_vs_fp64_oracle(slices):
    """G(250, 0.1), weights +-1..+-10: the root and three nodes with 5-20
    fixings, three outer rounds with triangle cuts, mixed products and
    certified bounds, against the FP64 oracle.  No node may need the FP64
    re-solve here."""
    n0 = 250
    SP = O.build_L(I.gnp_adj(n0, 0.1, 3, weights="int10"))
    P = api.make_params(include_Pent=0, include_Hepta=0, max_outer_iter=3, extra_iter=0)
    xf, sx = I.random_fixings(n0, 20, 4, 9)
    xf[0] = 0
    sx[0] = 0
    seeds = np.arange(5, 9, dtype=np.uint64)
    r, outs = _bound_both(SP, P, xf, sx, seeds, slices)
    for b, o in enumerate(outs):
        print(f"node {b}: status {r['status'][b]} / oracle {o.status}", flush=True)
        assert r["status"][b] == 0 and o.status == 0
        rel = (r["bound"][b] - o.bound) / abs(o.bound)
        print(f"node {b}: n {r['stats'][b][7]}, mixed certified {r['bound'][b]:.6f}, FP64 oracle {o.bound:.6f}, "
              f"rel {rel:+.2e}", flush=True)
        assert abs(rel) <= 1e-4
        assert r["bound"][b] >= r["LB"][b]
        assert not r["stats"][b][8] & api.FLAG_FP64_RETRY


def test_certified_fp64_bound_tracks_plain_bound():
    """FP64 products with certification: same trajectory as the plain
    bound; each certified bound at most 1e-7 relative above it."""
    n0 = 60
    SP = O.build_L(I.gnp_adj(n0, 0.5, 1, weights="unit"))
    P = api.make_params(include_Pent=0, include_Hepta=0, max_outer_iter=3, extra_iter=0)
    xf, sx = I.random_fixings(n0, 10, 6, 2)
    seeds = np.arange(1, 7, dtype=np.uint64)
    ctx = api.BoundContext(SP, params=P, max_nodes=6)
    plain = ctx.bound(xf, sx, LB=0.0, diff=0.0, is_root=0, seeds=seeds)
    ctx.set_precision(slices=0, certify=1)
    cert = ctx.bound(xf, sx, LB=0.0, diff=0.0, is_root=0, seeds=seeds)
    ctx.close()
    for b in range(6):
        assert plain["status"][b] == 0 and cert["status"][b] == 0
        d = cert["bound"][b] - plain["bound"][b]
        assert 0.0 <= d <= 1e-7 * abs(plain["bound"][b]), (b, cert["bound"][b], plain["bound"][b])


def test_certified_bounds_only_change_the_prune_tests():
    """Certification replaces the value the prune tests read and nothing
    else: the bundle's initial t (bounding.c:133), the give-up tests and
    the root's diff keep reading f + fixedvalue, so a root bound follows the
    plain trajectory bit for bit (every round's f), and a branch and bound
    visits the same nodes in the same launch rounds."""
    n0 = 60
    SP = api.maxcut_L(I.gnp_adj(n0, 0.5, 2, weights="unit"))
    P = api.make_params(include_Pent=0, include_Hepta=0)
    P.TriIneq = 10 * n0                       # src/process_input.c:135-136
    z = np.zeros(n0 - 1, np.int32)
    runs = []
    for cert in (0, 1):
        ctx = api.BoundContext(SP, params=P, max_nodes=1)
        ctx.set_precision(slices=0, certify=cert)
        ctx.capture(True)
        ctx.push(z, z, LB=0.0, diff=0.0, is_root=1, seed=2020, tag=0)
        while ctx.busy():
            ctx.advance()
        runs.append((ctx.pop(), ctx.trajectories()[0]))
        ctx.close()
    (r0, t0), (r1, t1) = runs
    assert r0["status"] == 0 and r1["status"] == 0 and t0["rounds"] > 5
    assert np.array_equal(t0["round_f"], t1["round_f"])
    assert np.array_equal(r0["stats"], r1["stats"]) and r0["diff"] == r1["diff"]
    assert 0.0 <= r1["bound"] - r0["bound"] <= 1e-7 * abs(r0["bound"])
    plain, _ = api.solve(SP, P, seed=2020, batch=64, warm=True)
    cert, _ = api.solve(SP, P, seed=2020, batch=64, warm=True, certify=1)
    assert plain.status == 0 and cert.status == 0 and cert.best == plain.best
    assert (cert.nodes, cert.batches) == (plain.nodes, plain.batches)


@pytest.mark.parametrize("name", ["g05_60.0", "g05_80.0"])
def test_mixed_certified_bab_proves_fixture_optimum(golden, name):
    path = os.path.join(golden, "rudy", name)
    SP = api.maxcut_L(api.read_rudy(path))
    P = api.make_params()
    P.TriIneq = 10 * SP.shape[0]
    st, sol = api.solve(SP, P, seed=2020, batch=32, slices=4)
    print(f"{name}: mixed certified B&B {st.nodes} nodes, {st.seconds:.2f} s", flush=True)
    assert st.status == 0
    assert st.best == fixture_value(path + "-expected_output")
    assert api.evaluate_solution(SP, sol) == st.best


@pytest.mark.parametrize("seed", [1, 2])
def test_mixed_certified_bab_brute_force(seed):
    adj = I.gnp_adj(20, 0.5, seed, weights="int10")
    SP = api.maxcut_L(adj)
    P = api.make_params()
    P.TriIneq = 200
    st, sol = api.solve(SP, P, seed=seed, batch=8, slices=4)
    assert st.status == 0 and st.best == brute_maxcut(adj)


def test_three_slices_fall_back_to_fp64_where_needed():
    """With 3 slices the products are accurate to about 1e-7 of the line
    maxima, which can stall a line search near the boundary of the cone
    (node 1 of this batch); such solves are repeated with FP64 products and
    flagged, and every node still ends with status 0 and a bound within
    1e-4 of the FP64 oracle."""
    n0 = 250
    SP = O.build_L(I.gnp_adj(n0, 0.1, 3, weights="int10"))
    P = api.make_params(include_Pent=0, include_Hepta=0, max_outer_iter=3, extra_iter=0)
    xf, sx = I.random_fixings(n0, 20, 4, 9)
    xf[0] = 0
    sx[0] = 0
    seeds = np.arange(5, 9, dtype=np.uint64)
    r, outs = _bound_both(SP, P, xf, sx, seeds, 3)
    for b, o in enumerate(outs):
        retried = bool(r["stats"][b][8] & api.FLAG_FP64_RETRY)
        print(f"node {b}: status {r['status'][b]}, FP64 re-solve {retried}", flush=True)
        assert r["status"][b] == 0 and o.status == 0
        assert abs(r["bound"][b] - o.bound) <= 1e-4 * abs(o.bound)


def test_automatic_precision_follows_the_class_table():
    """slices = -1 takes the context's class entry of the automatic table:
    at n0 = 60 FP64 products with certificates (bitwise the certified FP64
    context), at n0 = 220 the 32 x 7 register tile with 4 bf16 slices
    (bitwise the explicit form, and within 1e-4 of the FP64 oracle)."""
    table = api.ipm_auto_table()
    assert table[224] == (3207, 4)
    assert all(table[ld][1] == 0 for ld in (64, 80, 96, 112, 128, 160, 192, 256))
    P = api.make_params(include_Pent=0, include_Hepta=0, max_outer_iter=2, extra_iter=0)
    for n0, nodes, seed in ((60, 4, 3), (220, 2, 4)):
        SP = O.build_L(I.gnp_adj(n0, 0.1 if n0 > 100 else 0.5, seed, weights="int10"))
        P.TriIneq = 10 * n0
        xf, sx = I.random_fixings(n0, 5, nodes, seed + 1)
        seeds = np.arange(1, nodes + 1, dtype=np.uint64)
        ctx = api.BoundContext(SP, params=P, max_nodes=nodes)
        ctx.set_precision(slices=-1)
        auto = ctx.bound(xf, sx, LB=0.0, diff=0.0, is_root=0, seeds=seeds)
        ctx.close()
        tile, sl = table[224] if n0 > 192 else (api.N.lib().bqb_ipm_tile_for(n0), 0)
        api.N.lib().bqb_set_ipm_tile(tile)
        ctx = api.BoundContext(SP, params=P, max_nodes=nodes)
        api.N.lib().bqb_set_ipm_tile(0)
        ctx.set_precision(slices=sl, certify=1)
        expl = ctx.bound(xf, sx, LB=0.0, diff=0.0, is_root=0, seeds=seeds)
        ctx.close()
        assert np.all(auto["status"] == 0) and np.array_equal(auto["bound"], expl["bound"]), (n0, auto["bound"],
                                                                                            expl["bound"])
        if n0 > 192:
            d = {f: getattr(P, f) for f, _ in api.BiqBinParameters._fields_}
            for b in range(nodes):
                o = O.sdpbound(SP, xf[b], sx[b], 0.0, 0.0, 0, int(seeds[b]), O.default_params(**d))[0]
                assert abs(auto["bound"][b] - o.bound) <= 1e-4 * abs(o.bound), (b, auto["bound"][b], o.bound)


def test_autotuner_for_automatic_precision():
    """The autotuner times the four variants of the classes it covers and
    installs valid entries (restored afterwards)."""
    default = api.ipm_auto_table()
    try:
        table, timed = api.autotune_ipm_auto(nmin=200, nmax=224, nodes=4, reps=1)
        assert set(timed) == {224}
        assert all(v > 0 for v in timed[224])
        assert table[224][0] in (3207, 22401) and table[224][1] in (0, 4)
        for ld in (64, 80, 96, 112, 128, 160, 192, 256):
            assert table[ld] == default[ld]
    finally:
        api.set_ipm_auto_table(default)
