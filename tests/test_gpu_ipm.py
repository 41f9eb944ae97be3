"""Parity of the batched HIP IPM (bqb_ipm_batch / ipm_mc_pk through the C ABI)
against the CPU oracle's restatement of src/ipm_mc_pk.c.

Tolerance: phi (the bound) within 1e-6 relative (the per-node bound
criterion); iteration counts and line-search step lengths identical; X within
1e-6 absolute."""
import numpy as np
import pytest

import oracle_py as O
from parallel_biqbin_maxcut_amd import api, instances as I

pytestmark = pytest.mark.gpu


def sub_matrices(n0, p, weights, seed, count, nfix_max):
    adj = I.gnp_adj(n0, p, seed, weights=weights)
    L = O.build_L(adj)
    xf, sx = I.random_fixings(n0, nfix_max, count, seed + 1)
    return [O.create_subproblem(L, xf[b], sx[b])[0] for b in range(count)]


def compare(Ls, r, trace=False):
    for b, L in enumerate(Ls):
        X, y, phi, st, tr = O.ipm(L, trace_cap=400 if trace else 0)
        assert r["status"][b] == 0, (b, r["status"][b])
        assert st.status == 0
        assert abs(r["phi"][b] - phi) <= 1e-6 * max(1.0, abs(phi)), (b, L.shape, r["phi"][b], phi)
        assert r["iters"][b] == st.iters, (b, r["iters"][b], st.iters)
        assert np.max(np.abs(r["X"][b] - X)) < 1e-6
        if trace:
            assert np.allclose(r["trace"][b], tr, rtol=0, atol=1e-12)


def test_ipm_batch_cfg2_like_ragged():
    """G(100, 0.3) +-1 weights, 0-40 fixings: sizes 60-100 in one launch."""
    Ls = sub_matrices(100, 0.3, "pm1", 3, 24, 40)
    r = api.ipm_batch(Ls, trace_cap=200)
    compare(Ls, r, trace=True)


def test_ipm_batch_unit_weights_cfg1_root():
    adj = I.gnp_adj(60, 0.5, 1, weights="unit")
    L, _ = O.create_subproblem(O.build_L(adj), np.zeros(59, np.int32), np.zeros(59, np.int32))
    r = api.ipm_batch([L], trace_cap=200)
    compare([L], r, trace=True)


@pytest.mark.parametrize("n0", [2, 3, 17, 64, 65, 97, 128, 151, 161, 200, 250])
def test_ipm_batch_sizes_and_tiles(n0):
    Ls = sub_matrices(n0, 0.5, "int10", n0, 3, min(5, n0 - 1))
    compare(Ls, api.ipm_batch(Ls))


def test_ipm_batch_cfg3_like():
    """G(250, 0.1), weights +-1..+-10: sizes 210-250 (tiles 32 x 7, 32 x 8)."""
    Ls = sub_matrices(250, 0.1, "int10", 31, 3, 40)
    compare(Ls, api.ipm_batch(Ls))


def test_ipm_batch_qubo_cfg4_like():
    Q = I.random_qubo(150, 0.5, 5)
    L = O.build_L(I.qubo_to_maxcut_adj(Q))
    xf, sx = I.random_fixings(151, 40, 6, 9)
    Ls = [O.create_subproblem(L, xf[b], sx[b])[0] for b in range(6)]
    compare(Ls, api.ipm_batch(Ls))


def test_dropin_ipm_mc_pk_matches_batch():
    Ls = sub_matrices(50, 0.4, "pm1", 8, 1, 0)
    X, phi = api.ipm_mc_pk(Ls[0])
    _, _, phio, _, _ = O.ipm(Ls[0])
    assert abs(phi - phio) <= 1e-6 * abs(phio)
    assert np.allclose(np.diag(X), 1.0)


def test_ipm_batch_closed_form_complete_graph():
    n = 40
    adj = np.ones((n, n)) - np.eye(n)
    Lap = np.diag(adj.sum(1)) - adj
    r = api.ipm_batch([Lap / 4])
    assert n * n / 4 - 1e-9 <= r["phi"][0] <= n * n / 4 + 1e-2


def _numpy_first_iteration(L):
    """One predictor-corrector step of src/ipm_mc_pk.c in numpy (dense solves)."""
    n = L.shape[0]
    y = np.abs(L).sum(0) + 1.1
    X = np.eye(n)
    Z = np.diag(y) - L
    mu = np.sum(Z * X) / (2 * n)
    Zi = np.linalg.inv(Z)
    M = Zi * X
    dy1 = np.linalg.solve(M, -np.ones(n))
    dX1 = Zi @ (-np.diag(dy1) @ X) - X
    dX1 = 0.5 * (dX1 + dX1.T)
    rhs = mu * np.diag(Zi) - (Zi * dX1) @ dy1
    dy2 = np.linalg.solve(M, rhs)
    dX = Zi @ (-np.diag(dy2) @ X - np.diag(dy1) @ dX1) + mu * Zi + dX1
    dX = 0.5 * (dX + dX.T)
    return Zi, np.linalg.inv(M), dX, dy1 + dy2


@pytest.mark.parametrize("n0", [20, 70, 100])
def test_ipm_first_iteration_internals(n0):
    L = sub_matrices(n0, 0.4, "pm1", 5, 1, 0)[0]
    Zi, Mi, dX, dy = _numpy_first_iteration(L)
    r = api.ipm_debug(L, 1)
    assert r["iters"] == 1
    y0 = np.abs(L).sum(0) + 1.1
    assert np.allclose(r["dgL"], np.diag(L))
    assert np.allclose(r["y0"], y0, rtol=1e-14), (r["y0"][:6], y0[:6])
    assert np.allclose(r["zd0"], y0 - np.diag(L), rtol=1e-14)
    assert np.allclose(r["scal"], [y0.sum(), np.trace(L), np.sum(y0 - np.diag(L)) / (2 * L.shape[0])])
    assert np.max(np.abs(r["Zi"] - Zi)) < 1e-10 * np.max(np.abs(Zi))
    assert np.max(np.abs(r["Mi"] - Mi)) < 1e-8 * np.max(np.abs(Mi))
    assert np.max(np.abs(r["dX"] - dX)) < 1e-8 * np.max(np.abs(dX))
    _, _, _, _, tr = O.ipm(L, trace_cap=400)
    assert np.allclose(r["trace"][:2], tr[:2], atol=1e-12)
