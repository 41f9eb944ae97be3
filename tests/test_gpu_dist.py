"""Unit parity of the workgroup-distributed FP64 primitives (bqb_dist.h)
against dense numpy: sweep inverse, lower Cholesky, the Cholesky
positive-definiteness test and symmetrisation, for each register tile."""
import numpy as np
import pytest

from parallel_biqbin_maxcut_amd import api

pytestmark = pytest.mark.gpu

TILES = [(1604, 64), (1607, 112), (3204, 128)]


def spd(n, seed):
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((n, n))
    return B @ B.T + n * np.eye(n)


@pytest.mark.parametrize("tile,nmax", TILES)
@pytest.mark.parametrize("frac", [0.02, 0.3, 0.77, 1.0])
def test_sweep_inverse(tile, nmax, frac):
    n = max(1, int(round(frac * nmax)))
    A = spd(n, n)
    out, ok = api.dist_selftest(tile, 0, A)
    assert ok == 1
    assert np.max(np.abs(out - np.linalg.inv(A))) < 1e-12 * np.max(np.abs(np.linalg.inv(A))) * n


@pytest.mark.parametrize("tile,nmax", TILES)
@pytest.mark.parametrize("frac", [0.02, 0.5, 1.0])
def test_chol_lower(tile, nmax, frac):
    n = max(1, int(round(frac * nmax)))
    A = spd(n, n + 1)
    out, ok = api.dist_selftest(tile, 1, A)
    assert ok == 1
    assert np.allclose(out, np.linalg.cholesky(A), atol=1e-10)


@pytest.mark.parametrize("tile,nmax", TILES)
def test_chol_test_detects_indefinite(tile, nmax):
    n = nmax - 3
    A = spd(n, 3)
    assert api.dist_selftest(tile, 2, A)[1] == 1
    w, V = np.linalg.eigh(A)
    w[n // 2] = -1e-3
    assert api.dist_selftest(tile, 2, (V * w) @ V.T)[1] == 0


@pytest.mark.parametrize("tile,nmax", TILES)
@pytest.mark.parametrize("n_off", [1, 2, 15, 16, 17, 31, 33, 48, 0])
def test_chol_test_first_failing_pivot(tile, nmax, n_off):
    """A matrix whose leading k x k blocks are PD up to pivot p and fail at
    p: the test must fail, whichever block (and position in it) p falls in;
    the same matrix cut to its first p rows must pass."""
    n = nmax - 1 if n_off == 0 else min(nmax, n_off + 20)
    p = n - 1 if n_off == 0 else min(n_off, n - 1)
    rng = np.random.default_rng(100 + n_off)
    L = np.tril(rng.standard_normal((n, n)) * 0.3) + np.diag(1.0 + rng.random(n))
    D = np.ones(n)
    D[p] = -0.5                      # pivot p of A = L D L' is negative
    A = (L * D) @ L.T
    assert api.dist_selftest(tile, 2, A)[1] == 0
    assert api.dist_selftest(tile, 2, A[:p, :p])[1] == 1 if p > 0 else True


@pytest.mark.parametrize("tile,nmax", TILES)
def test_symmetrize(tile, nmax):
    n = nmax - 1
    A = np.random.default_rng(1).standard_normal((n, n))
    out, _ = api.dist_selftest(tile, 3, A)
    assert np.array_equal(out, 0.5 * (A + A.T))
