"""Pins of the oracle's bundle method and triangle update by what the
mathematics and the reference's stated rules fix, and a demonstration that
each check notices a changed constant.

* lambda_eta (bundle.c:292-373): its output gamma + dgamma must solve the
  proximal bundle subproblem
      min_{g >= 0}  max_j (F_j + G_j'g) + |g - gamma|^2 / (2t),
  which is solved here independently (scipy); eta and dgamma must satisfy
  eta >= 0, gamma + dgamma >= 0, eta o (gamma + dgamma) = 0 and
  dgamma = t (eta - G lambda); lambda lies on the simplex; t leaves
  unchanged, or multiplied by 0.95 when the loop stops at 50 rounds.
* bundle_method (bundle.c:86-233) along whole node bounds: every step is
  serious exactly when f - f_test > 0.05 del; a serious step sets f = f_test
  and t *= 1.01, a null step keeps f and divides t by 1.01; the purge drops
  exactly the elements with lambda_i < 0.01 max(lambda); after a cut change t
  is multiplied by 1.05 (bounding.c:318).
* updateTriangleInequalities (cutting_planes.c:132-214): the survivors are
  the cuts with y >= 1e-5 in their order, with their y; the separated list is
  a brute-force replay of the replace-the-least rule of cutting_planes.c:48-127
  (set, order and values); the appended cuts are the list minus the
  survivors, in list order, with y = 0.

test_checks_detect_changed_constants builds the oracle with one constant
changed at a time and asserts that some check above fails for each."""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest
from scipy.optimize import minimize

import oracle_py as O
from parallel_biqbin_maxcut_amd import instances as I

ORACLE_DIR = os.path.dirname(os.path.abspath(O.__file__))


# ------------------------------------------------------------ lambda_eta
def _prox_objective(gp, F, G, gamma, t):
    return np.max(F + G.T @ gp) + np.sum((gp - gamma) ** 2) / (2 * t)


def _prox_solve(F, G, gamma, t):
    """The proximal bundle subproblem by SLSQP in epigraph form."""
    m, k = G.shape
    x0 = np.concatenate([np.maximum(gamma, 0), [np.max(F + G.T @ gamma)]])
    cons = [{"type": "ineq", "fun": lambda x, j=j: x[m] - F[j] - G[:, j] @ x[:m],
             "jac": lambda x, j=j: np.concatenate([-G[:, j], [1.0]])} for j in range(k)]
    res = minimize(lambda x: x[m] + np.sum((x[:m] - gamma) ** 2) / (2 * t), x0,
                   jac=lambda x: np.concatenate([(x[:m] - gamma) / t, [1.0]]),
                   bounds=[(0, None)] * m + [(None, None)], constraints=cons, method="SLSQP",
                   options={"ftol": 1e-14, "maxiter": 2000})
    gp = np.maximum(res.x[:m], 0)
    return gp, _prox_objective(gp, F, G, gamma, t)


def _lambda_eta_case(seed):
    rng = np.random.default_rng(seed)
    m, k = int(rng.integers(2, 40)), int(rng.integers(2, 16))
    G = rng.normal(size=(m, k)) * rng.choice([0.3, 1, 3])
    F = rng.normal(size=k) * rng.choice([0.3, 1, 3])
    gamma = np.maximum(rng.normal(size=m), 0) * rng.choice([0.1, 1])
    t = float(10 ** rng.uniform(-2, 1))
    return F, G, gamma, t


def check_lambda_eta_identities(seeds=range(40)):
    for s in seeds:
        F, G, gamma, t = _lambda_eta_case(s)
        dg, lam, eta, t_out, st = O.lambda_eta(-F - G.T @ gamma, G, gamma, t)
        assert st == 0
        gp = gamma + dg
        assert np.all(eta >= 0)
        assert np.all(gp >= -1e-12)
        assert np.max(np.abs(eta * gp)) <= 1e-10 * (1 + np.max(np.abs(eta)))
        assert np.allclose(dg, t * (eta - G @ lam), rtol=0, atol=1e-11 * (1 + np.max(np.abs(dg))))
        assert np.all(lam >= 0) and abs(lam.sum() - 1) <= 1e-9
        assert t_out == t or t_out == t * 0.95, (s, t, t_out)


def check_lambda_eta_optimal(seeds=range(40)):
    """The proximal point is optimal to the accuracy that the reference's
    stop rule gives (it stops when |dgamma| changes by less than 1e-5
    relative between two lambda-eta rounds, which leaves the objective up to
    about 2e-3 relative above the optimum on these instances; a stop rule of
    1e-2 leaves up to 0.4)."""
    for s in seeds:
        F, G, gamma, t = _lambda_eta_case(s)
        dg, lam, eta, t_out, st = O.lambda_eta(-F - G.T @ gamma, G, gamma, t)
        if t_out != t:
            continue                          # stopped at the 50-round cap with t changed
        _, best = _prox_solve(F, G, gamma, t)
        got = _prox_objective(gamma + dg, F, G, gamma, t)
        assert got >= best - 1e-7 * (1 + abs(best))
        assert got - best <= 5e-3 * (1 + abs(best)), (s, got, best)


def check_lambda_eta_cap():
    """Instances whose lambda-eta loop runs into the 50-round cap (found by
    search): t must come back multiplied by 0.95."""
    capped = 0
    for s in range(120):
        rng = np.random.default_rng(s)
        m, k = int(rng.integers(2, 60)), int(rng.integers(2, 20))
        G = rng.normal(size=(m, k)) * rng.choice([0.1, 1, 10])
        F = rng.normal(size=k) * rng.choice([0.1, 1, 10])
        gamma = np.maximum(rng.normal(size=m), 0) * rng.choice([0.1, 1, 10])
        t = float(10 ** rng.uniform(-3, 3))
        _, _, _, t_out, st = O.lambda_eta(-F - G.T @ gamma, G, gamma, t)
        assert st == 0
        assert t_out == t or t_out == t * 0.95, (s, t, t_out)
        capped += t_out != t
    assert capped >= 3


def test_lambda_eta_identities():
    check_lambda_eta_identities()


def test_lambda_eta_solves_the_proximal_subproblem():
    check_lambda_eta_optimal()


def test_lambda_eta_iteration_cap_shrinks_t():
    check_lambda_eta_cap()


# ------------------------------------------------------------ bundle law
def _node_traces():
    """Step traces of whole node bounds: a root and a node with fixings, with
    and without pentagonal/heptagonal cuts."""
    out = []
    for n0, p, w, seed, nfix, kgon in ((40, 0.5, "unit", 1, 0, 0), (50, 0.4, "pm1", 3, 6, 1),
                                       (45, 0.6, "int10", 5, 3, 1)):
        SP = O.build_L(I.gnp_adj(n0, p, seed, weights=w))
        P = O.reference_params_file(include_Pent=kgon, include_Hepta=kgon, TriIneq=10 * n0)
        xf, sx = I.random_fixings(n0, nfix, 1, seed)
        if nfix == 0:
            xf[:] = 0
            sx[:] = 0
        o, _, _, rows = O.sdpbound_trace(SP, xf[0], sx[0], 0.0, 0.0, 1, seed, P)
        assert o.status == 0
        out.append(rows)
    return out


def check_bundle_law(traces):
    steps = 0
    for rows in traces:
        t_prev_step = None
        rebuild_pending = False
        i = 0
        while i < len(rows):
            r = rows[i]
            if r[0] == 3:                      # end of an outer round: a cut change may follow
                rebuild_pending = True
            if r[0] == 2:                      # QP of a bundle step: t after lambda_eta
                t_qp = r[2]
                if t_prev_step is not None and rebuild_pending:
                    assert t_qp in (t_prev_step * 1.05, t_prev_step * 1.05 * 0.95), (t_prev_step, t_qp)
                elif t_prev_step is not None:
                    assert t_qp in (t_prev_step, t_prev_step * 0.95), (t_prev_step, t_qp)
                dec = rows[i + 1]
                assert dec[0] == 5
                _, cnt, dl, f_before, f_test, lmax, k, serious = dec
                k = int(k)
                lam = rows[i + 2: i + 2 + k]
                # rows 6: (6, count, i, lambda_i, max lambda, k, serious, 0)
                assert np.all(lam[:, 0] == 6) and np.array_equal(lam[:, 2], np.arange(k))
                step = rows[i + 2 + k]
                assert step[0] == 1
                rule = (f_before - f_test > 0.05 * dl)
                assert bool(serious) == rule, (f_before, f_test, dl)
                assert lmax == np.max(lam[:, 3]) and np.all(lam[:, 4] == lmax)
                purge_range = k if serious else k - 1
                kept = int(np.sum(lam[:purge_range, 3] >= 0.01 * lmax))
                assert int(step[5]) == kept + 1 + (0 if serious else 1), (k, kept, step[5])
                if serious:
                    assert step[3] == f_test and step[4] == t_qp * 1.01
                else:
                    assert step[3] == f_before and step[4] == t_qp / 1.01
                t_prev_step = step[4]
                rebuild_pending = False
                steps += 1
                i += 3 + k
                continue
            i += 1
    assert steps > 50


def test_bundle_steps_follow_the_serious_null_law():
    check_bundle_law(_node_traces())


# ------------------------------------------------------------ triangle update
def _tri_value(X, t, i, j, k):
    """Triangle inequality value of cutting_planes.c:20-39 for i > j > k."""
    a, b, c = X[i, j], X[i, k], X[j, k]
    return {1: -a - b - c - 1, 2: -a + b + c - 1, 3: a - b + c - 1, 4: a + b - c - 1}[t]


def _replay_list(X, threshold, cap):
    """Brute-force replay of the replace-the-least rule (cutting_planes.c:48-127)."""
    n = X.shape[0]
    lst, least, least_v = [], 0, None
    for t in range(1, 5):
        for i in range(n):
            for j in range(i):
                for k in range(j):
                    v = _tri_value(X, t, i, j, k)
                    if v <= threshold:
                        continue
                    if len(lst) < cap:
                        lst.append([i, j, k, t, v])
                        if len(lst) == 1 or v < least_v:
                            least_v, least = v, len(lst) - 1
                    elif v > least_v:
                        lst[least] = [i, j, k, t, v]
                        least_v = min(e[4] for e in lst)
                        least = [e[4] for e in lst].index(least_v)
    return [tuple(e) for e in lst]


def _rand_X(n, seed):
    """A correlation matrix of rank 3 (violates many triangle inequalities)."""
    rng = np.random.default_rng(seed)
    V = rng.normal(size=(n, 3))
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    return np.asfortranarray(V @ V.T)


def check_triangle_update(seeds=range(6)):
    for s in seeds:
        rng = np.random.default_rng(100 + s)
        n = int(rng.integers(7, 13))
        X = _rand_X(n, s)
        P = O.default_params(violated_TriIneq=0.05, TriIneq=int(rng.integers(5, 30)))
        full = _replay_list(X, 0.05, 10 ** 6)
        cur = [full[q][:4] for q in rng.permutation(len(full))[: min(len(full), 12)]]
        cur += [(4, 2, 1, 1), (5, 3, 0, 2)]                      # possibly not violated any more
        y = rng.choice([0.0, 5e-6, 2e-5, 5e-5, 0.3, 1.2], size=len(cur))
        out, lst, mx = O.update_triangles_on(X, P, cur, y, 10000)
        exp_list = _replay_list(X, 0.05, P.TriIneq)
        assert [e[:4] for e in lst] == [e[:4] for e in exp_list]
        assert np.allclose([e[4] for e in lst], [e[4] for e in exp_list], rtol=0, atol=1e-14)
        surv = [(c[0], c[1], c[2], c[3], yy) for c, yy in zip(cur, y) if yy >= 1e-5]
        assert out[: len(surv)] == surv
        skeys = {c[:4] for c in surv}
        app = [(e[0], e[1], e[2], e[3], 0.0) for e in exp_list if e[:4] not in skeys]
        assert out[len(surv):] == app
        assert mx == pytest.approx(max(_tri_value(X, t, i, j, k) for t in range(1, 5) for i in range(n)
                                       for j in range(i) for k in range(j)), abs=1e-14)


def test_triangle_update_matches_brute_force_replay():
    check_triangle_update()


# ------------------------------------------------------------ mutations
MUTANTS = {
    "ORC_SERIOUS_FRAC=0.1": "bundle",
    "ORC_T_STEP=1.02": "bundle",
    "ORC_PURGE_LAMBDA=0.02": "bundle",
    "ORC_T_REBUILD=1.1": "bundle",
    "ORC_T_NOCONV=0.9": "cap",
    "ORC_LE_TOL=1e-2": "optimal",
    "ORC_CUT_PURGE=1e-4": "triangle",
}


def _build_variant(define, dst):
    srcs = ["orc_ipm.c", "orc_cuts.c", "orc_bundle.c", "orc_heur.c", "orc_bound.c", "orc_many.c"]
    cmd = ["gcc", "-O1", "-fPIC", "-std=c11", "-D_POSIX_C_SOURCE=200809L", "-fopenmp", "-shared",
           "-D" + define, "-o", dst] + [os.path.join(ORACLE_DIR, s) for s in srcs] + ["-lm"]
    subprocess.run(cmd, check=True, capture_output=True)


def test_checks_detect_changed_constants():
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    from concurrent.futures import ThreadPoolExecutor
    tmp = tempfile.mkdtemp(prefix="orc_mut_")
    try:
        paths = {d: os.path.join(tmp, f"lib_{i}.so") for i, d in enumerate(MUTANTS)}
        with ThreadPoolExecutor(max_workers=4) as ex:
            list(ex.map(lambda d: _build_variant(d, paths[d]), MUTANTS))
        checks = {"bundle": lambda: check_bundle_law(_node_traces()),
                  "cap": check_lambda_eta_cap,
                  "optimal": check_lambda_eta_optimal,
                  "triangle": check_triangle_update}
        for define, which in MUTANTS.items():
            with O.variant(paths[define]):
                with pytest.raises(AssertionError):
                    checks[which]()
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
