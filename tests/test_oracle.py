"""Pins of the CPU oracle: closed forms, invariants, brute force, and the
reference's own end-to-end fixtures (tests/rudy, tests/qubos of the reference).

These run on the CPU only; they establish that the oracle is a faithful
checker before it is used to judge the HIP path."""
import itertools
import json
import os

import numpy as np
import pytest

import oracle_py as O
from fixture_util import brute_maxcut, fixture_solution, fixture_value, laplacian
from parallel_biqbin_maxcut_amd import instances as I


# ---------------------------------------------------------------- IPM
@pytest.mark.parametrize("n", [3, 6, 9, 12])
def test_ipm_complete_graph_closed_form(n):
    """K_n with unit weights: max <L/4, X> s.t. diag X = e is n^2/4 (X = (nI - ee')/(n-1))."""
    adj = np.ones((n, n)) - np.eye(n)
    X, y, phi, st, _ = O.ipm(laplacian(adj) / 4)
    assert st.status == 0
    assert n * n / 4 - 1e-9 <= phi <= n * n / 4 + 1e-2


def test_ipm_five_cycle_closed_form():
    """C5: SDP value (5/2)(1 + cos(pi/5)) (Goemans-Williamson)."""
    adj = np.zeros((5, 5))
    for i in range(5):
        adj[i, (i + 1) % 5] = adj[(i + 1) % 5, i] = 1
    X, y, phi, st, _ = O.ipm(laplacian(adj) / 4)
    exact = 2.5 * (1 + np.cos(np.pi / 5))
    assert exact - 1e-9 <= phi <= exact + 1e-2


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_ipm_invariants_and_weak_duality(seed):
    n = 12
    adj = I.gnp_adj(n, 0.5, seed, weights="pm1")
    L = laplacian(adj) / 4
    X, y, phi, st, trace = O.ipm(L, trace_cap=200)
    psi = float(np.sum(L * X))
    assert st.status == 0 and st.iters > 0
    assert abs(phi - psi) <= 1e-2                      # exit test (ipm_mc_pk.c:126)
    assert np.allclose(np.diag(X), 1.0, atol=1e-12)    # X = I + sum of symmetric steps with zero diagonal drift
    assert np.linalg.eigvalsh(X).min() > 0
    Z = np.diag(y) - L
    # dual feasibility of the returned y: Diag(y) - L PSD (the bound is valid at any gap)
    assert np.linalg.eigvalsh(Z).min() > -1e-9
    assert abs(phi - y.sum()) < 1e-9
    assert phi >= brute_maxcut(adj) - 1e-9
    assert len(trace) == 2 * st.iters
    assert np.all((trace > 0) & (trace <= 1))


def test_ipm_first_alpha_steps_are_backtracking_powers():
    """alpha in {1} U {0.95 * 0.8^j}: the line-search law of ipm_mc_pk.c:281-309."""
    adj = I.gnp_adj(20, 0.3, 7, weights="int10")
    _, _, _, st, trace = O.ipm(laplacian(adj) / 4, trace_cap=400)
    for a in trace:
        if a != 1.0:
            j = np.log(a / 0.95) / np.log(0.8)
            assert abs(j - round(j)) < 1e-9


# ---------------------------------------------------------------- cut operators
def random_cuts(n, rng, nt=6, np_=4, nh=3):
    tris = []
    for _ in range(nt):
        i, j, k = sorted(rng.choice(n, 3, replace=False))[::-1]
        tris.append((int(i), int(j), int(k), int(rng.integers(1, 5))))
    pents = [(int(rng.integers(1, 4)), [int(v) for v in rng.choice(n, 5, replace=False)]) for _ in range(np_)]
    hepts = [(int(rng.integers(1, 5)), [int(v) for v in rng.choice(n, 7, replace=False)]) for _ in range(nh)]
    return tris, pents, hepts


def test_op_B_op_Bt_adjoint():
    rng = np.random.default_rng(0)
    n = 11
    tris, pents, hepts = random_cuts(n, rng)
    cuts = O.cut_arrays(tris, pents, hepts)
    A = rng.standard_normal((n, n))
    X = A + A.T
    t = rng.standard_normal(len(tris) + len(pents) + len(hepts))
    BX = -O.op_B(X, cuts)                            # op_B computes y -= B(X)
    BtT = -(O.op_Bt(np.zeros((n, n)), cuts, t))      # op_Bt computes X -= B'(t)
    assert abs(BX @ t - np.sum(X * BtT)) < 1e-10
    assert np.allclose(BtT, BtT.T) and np.all(np.diag(BtT) == 0)


@pytest.mark.parametrize("k,types", [(3, [1, 2, 3, 4]), (5, [1, 2, 3]), (7, [1, 2, 3, 4])])
def test_cuts_valid_and_tight_for_all_cuts(k, types):
    """Every cut B_c(X) <= 1 holds for all X = xx', x in {-1,1}^n, and is tight."""
    n = k
    for t in types:
        if k == 3:
            cuts = O.cut_arrays([(2, 1, 0, t)])
        elif k == 5:
            cuts = O.cut_arrays(pents=[(t, list(range(5)))])
        else:
            cuts = O.cut_arrays(hepts=[(t, list(range(7)))])
        vals = []
        for s in itertools.product([-1, 1], repeat=n):
            x = np.array(s, dtype=float)
            vals.append(-O.op_B(np.outer(x, x), cuts)[0])
        assert max(vals) == pytest.approx(1.0)


def test_triangle_types_are_the_four_facets():
    """The four triangle types are the four distinct sign patterns with an even
    number of minus signs flipped: each is violated by some X with |X_ij|<=1."""
    seen = set()
    for t in (1, 2, 3, 4):
        cuts = O.cut_arrays([(2, 1, 0, t)])
        coef = []
        for (a, b) in [(2, 1), (2, 0), (1, 0)]:
            E = np.zeros((3, 3)); E[a, b] = E[b, a] = 1
            coef.append(int(round(-O.op_B(E, cuts)[0])))
        seen.add(tuple(coef))
        assert int(np.prod(coef)) == -1          # odd number of minus signs
    assert len(seen) == 4


def test_sa_cost_matches_op_B_and_H_tables():
    """cost = <H, X(p,p)> and op_B agree: B_c(X) = (k - cost)/(2(k-1)/ ... )."""
    rng = np.random.default_rng(3)
    n = 14
    A = rng.standard_normal((n, 6))
    X = A @ A.T
    d = np.sqrt(np.diag(X)); X = X / np.outer(d, d)
    Xf = O.f64(X)
    for k, ntypes, coef in [(5, 3, 0.5), (7, 4, 1.0 / 3.0)]:
        for t in range(1, ntypes + 1):
            Hp = O.lib().orc_H_pent(t) if k == 5 else O.lib().orc_H_hept(t)
            H = np.array([Hp[i] for i in range(k * k)]).reshape(k, k)
            e = np.array([-1 if a < t - 1 else 1 for a in range(k)])
            assert np.array_equal(H, np.outer(e, e))    # H = e e' with the first t-1 entries of e negative
            perm = (O.C.c_int * n)()
            st = O.C.c_int(0)
            cost = O.lib().orc_qap_sa(O._i(np.ascontiguousarray(H.ravel(), dtype=np.int32)), k,
                                      O._d(Xf), n, perm, 12345 + t, O.C.byref(st))
            p = [perm[i] for i in range(k)]
            assert len(set(p)) == k and st.value == 0
            assert cost == pytest.approx(float(np.sum(H * X[np.ix_(p, p)])), abs=1e-9)
            cuts = O.cut_arrays(pents=[(t, p)]) if k == 5 else O.cut_arrays(hepts=[(t, p)])
            B = -O.op_B(X, cuts)[0]
            assert B == pytest.approx(coef * (k - cost) / 2.0, abs=1e-9)


def test_triangle_scan_topk_and_values():
    rng = np.random.default_rng(5)
    n = 16
    A = rng.standard_normal((n, 3))
    X = A @ A.T
    d = np.sqrt(np.diag(X)); X = X / np.outer(d, d)
    thr, cap = 0.01, 25
    mx, lst = O.triangles(X, thr, cap)
    allv = []
    for t in (1, 2, 3, 4):
        for i in range(n):
            for j in range(i):
                for k in range(j):
                    cuts = O.cut_arrays([(i, j, k, t)])
                    allv.append(-O.op_B(X, cuts)[0] - 1.0)
    allv = np.array(allv)
    assert mx == pytest.approx(allv.max(), abs=1e-12)
    viol = np.sort(allv[allv > thr])[::-1]
    got = np.sort([c[4] for c in lst])[::-1]
    assert len(got) == min(cap, len(viol))
    assert np.allclose(got, viol[: len(got)], atol=1e-12)
    for (i, j, k, t, v) in lst:
        assert i > j > k
        assert v == pytest.approx(-O.op_B(X, O.cut_arrays([(i, j, k, t)]))[0] - 1.0, abs=1e-12)


# ---------------------------------------------------------------- heuristics
def test_mc_1opt_local_optimum():
    rng = np.random.default_rng(9)
    n = 30
    adj = I.gnp_adj(n, 0.4, 11, weights="pm1")
    L = laplacian(adj) / 4
    Lf = O.f64(L)
    for _ in range(5):
        x = np.ascontiguousarray(rng.choice([-1, 1], n), dtype=np.int32)
        cost = O.lib().orc_mc_1opt(O._i(x), O._d(Lf), n)
        assert cost == pytest.approx(x @ L @ x)
        for i in range(n):
            y = x.copy(); y[i] *= -1
            assert y @ L @ y <= x @ L @ x + 0.001 * 4 + 1e-12


# ---------------------------------------------------------------- bundle QP
@pytest.mark.parametrize("k", [1, 2, 5, 12])
def test_solve_lambda_kkt(k):
    rng = np.random.default_rng(k)
    G = rng.standard_normal((8, k))
    Q = 0.7 * G.T @ G
    c = rng.standard_normal(k)
    lam = np.zeros(k)
    st = O.lib().orc_solve_lambda(k, O._d(O.f64(Q)), O._d(c), O._d(lam))
    assert st == 0
    assert np.all(lam >= 0) and abs(lam.sum() - 1) < 1e-8
    obj = lambda l: l @ c + 0.5 * l @ Q @ l
    # compare with a dense search over random simplex points refined near vertices
    best = obj(lam)
    for _ in range(2000):
        v = rng.dirichlet(np.ones(k) * 0.3)
        assert obj(v) >= best - 1e-4


# ---------------------------------------------------------------- subproblem
def test_create_subproblem_brute_force():
    rng = np.random.default_rng(2)
    n0 = 9
    adj = I.gnp_adj(n0, 0.6, 21, weights="int10")
    SP = O.build_L(adj)
    B = n0 - 1
    for trial in range(6):
        xf = np.zeros(B, dtype=np.int32); sx = np.zeros(B, dtype=np.int32)
        fix = rng.choice(B, int(rng.integers(0, 4)), replace=False)
        xf[fix] = 1; sx[fix] = rng.integers(0, 2, len(fix))
        PL, fv = O.create_subproblem(SP, xf, sx)
        free = [i for i in range(B) if not xf[i]]
        best_direct, best_sub = -1e18, -1e18
        for bits in range(1 << len(free)):
            x = sx.copy()
            for b, i in enumerate(free):
                x[i] = (bits >> b) & 1
            best_direct = max(best_direct, O.evaluate_solution(SP, x))
            s = np.array([2 * x[i] - 1 for i in free] + [-1.0])
            best_sub = max(best_sub, s @ PL @ s + fv)
        assert best_sub == pytest.approx(best_direct)


def test_fct_eval_weak_duality():
    rng = np.random.default_rng(4)
    n0 = 10
    adj = I.gnp_adj(n0, 0.5, 5, weights="pm1")
    SP = O.build_L(adj)
    PL, fv = O.create_subproblem(SP, np.zeros(n0 - 1, dtype=np.int32), np.zeros(n0 - 1, dtype=np.int32))
    exact = brute_maxcut(adj)
    n = PL.shape[0]
    tris, pents, hepts = random_cuts(n, rng, 10, 3, 2)
    cuts = O.cut_arrays(tris, pents, hepts)
    f0, X0, g0, _ = O.fct_eval(PL, cuts, np.zeros(15))
    _, _, phi, _, _ = O.ipm(PL)
    assert f0 == pytest.approx(phi, abs=1e-12)                    # f(0) = phi(L)
    assert np.allclose(g0, 1.0 + O.op_B(X0, cuts), atol=1e-12)    # g = 1 - B(X)
    for _ in range(4):
        gam = rng.uniform(0, 0.3, 15)
        f, X, g, st = O.fct_eval(PL, cuts, gam)
        assert st.status == 0 and f + fv >= exact - 1e-9


# ---------------------------------------------------------------- end to end
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_bab_small_graphs_brute_force(seed):
    adj = I.gnp_adj(14, 0.5, seed, weights="int10")
    out, sol = O.bab(O.build_L(adj), O.reference_params_file(), seed=seed)
    assert out.status == 0
    assert out.best == pytest.approx(brute_maxcut(adj))
    assert O.evaluate_solution(O.build_L(adj), sol) == pytest.approx(out.best)


def _rudy_case(golden, name, slow_ok=True):
    adj = I.read_rudy(os.path.join(golden, "rudy", name))
    L = O.build_L(adj)
    exp_path = os.path.join(golden, "rudy", name + "-expected_output")
    out, sol = O.bab(L, O.reference_params_file(), seed=1)
    assert out.status == 0
    assert out.best == fixture_value(exp_path)
    fsol = np.zeros(len(sol), dtype=np.int32)
    fsol[np.array(fixture_solution(exp_path)) - 1] = 1
    assert O.evaluate_solution(L, fsol) == out.best                # the fixture's vector attains it
    assert O.evaluate_solution(L, sol) == out.best
    return out, sol, fsol


def test_fixture_g05_60(golden):
    out, sol, fsol = _rudy_case(golden, "g05_60.0")
    assert out.nodes == 1
    assert round(out.root_bound, 2) == 536.95                     # g05_60.0-expected_output:27


@pytest.mark.slow
def test_fixture_g05_100(golden):
    out, sol, fsol = _rudy_case(golden, "g05_100.4")
    assert np.array_equal(sol, fsol)                               # same optimal vector


@pytest.mark.slow
def test_fixture_g05_80(golden):
    _rudy_case(golden, "g05_80.0")


def test_fixture_kcluster40_qubo(golden):
    Q = I.read_sparse_json(os.path.join(golden, "qubos", "kcluster40_025_10_1.json"), "qubo")
    exp = json.load(open(os.path.join(golden, "qubos", "kcluster40_025_10_1.json-expected_output.json")))
    L = O.build_L(I.qubo_to_maxcut_adj(Q))
    out, sol = O.bab(L, O.reference_params_file(), seed=1)
    assert out.best == pytest.approx(exp["maxcut"]["computed_val"], abs=1e-4)
    y = I.maxcut_solution_to_qubo([i + 1 for i in range(len(sol)) if sol[i]], Q.shape[0])
    assert y @ Q @ y == pytest.approx(exp["qubo"]["computed_val"], abs=1e-4)
