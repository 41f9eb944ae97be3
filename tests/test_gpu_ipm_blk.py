"""The blocked IPM (tile codes ld*100 + 1, selected for large subproblems)
against the CPU oracle's restatement of src/ipm_mc_pk.c, with the same
criteria as the register-tile kernel: phi within 1e-6 relative, the same
iteration count, step lengths within 1e-12, X within 1e-6; and the
first-iteration internals (Zi, Mi, dX) against the oracle's."""
import numpy as np
import pytest

import oracle_py as O
from parallel_biqbin_maxcut_amd import api, instances as I
from test_gpu_ipm import compare, sub_matrices

pytestmark = pytest.mark.gpu

LDS = (64, 80, 96, 112, 128, 160, 192, 224, 256)


def blocked_code(nmax):
    return next(ld for ld in LDS if ld >= nmax) * 100 + 1


class forced:
    def __init__(self, tile):
        self.tile = tile

    def __enter__(self):
        assert api.N.lib().bqb_set_ipm_tile(self.tile) == 0

    def __exit__(self, *a):
        api.N.lib().bqb_set_ipm_tile(0)


@pytest.mark.parametrize("n0", [2, 3, 17, 40, 64, 65, 100, 129, 151, 161, 200, 241, 250])
def test_blocked_ipm_sizes(n0):
    Ls = sub_matrices(n0, 0.5, "int10", n0, 3, min(5, n0 - 1))
    with forced(blocked_code(n0)):
        r = api.ipm_batch(Ls, trace_cap=400)
    compare(Ls, r, trace=True)


def test_blocked_ipm_cfg3_like_ragged():
    """G(250, 0.1), weights +-1..+-10, 0-40 fixings: sizes 210-250."""
    Ls = sub_matrices(250, 0.1, "int10", 31, 6, 40)
    with forced(25601):
        r = api.ipm_batch(Ls, trace_cap=400)
    compare(Ls, r, trace=True)


def test_blocked_ipm_cfg4_like_qubo():
    Q = I.random_qubo(150, 0.5, 7)
    Ls = sub_matrices_from_L(O.build_L(I.qubo_to_maxcut_adj(Q)), 4, 40, 8)
    with forced(16001):
        r = api.ipm_batch(Ls, trace_cap=400)
    compare(Ls, r, trace=True)


def sub_matrices_from_L(L, count, nfix, seed):
    n0 = L.shape[0]
    xf, sx = I.random_fixings(n0, nfix, count, seed)
    return [O.create_subproblem(L, xf[b], sx[b])[0] for b in range(count)]


@pytest.mark.parametrize("n0", [70, 180])
def test_blocked_first_iteration_internals(n0):
    """After one iteration: Zi = inv(Z0), Mi = inv(Zi o X0), dX against the
    register-tile kernel's (both within 1e-9 of each other)."""
    Ls = sub_matrices(n0, 0.5, "int10", 5, 1, 3)
    L = Ls[0]
    reg = api.ipm_debug(L, 1)
    with forced(blocked_code(L.shape[0])):
        blk = api.ipm_debug(L, 1)
    for k in ("Zi", "Mi", "dX", "X", "y"):
        a, b = blk[k], reg[k]
        assert np.max(np.abs(a - b)) <= 1e-9 * max(1.0, np.max(np.abs(b))), k
    assert np.array_equal(blk["y0"], reg["y0"]) and np.array_equal(blk["dgL"], reg["dgL"])


def test_blocked_ipm_in_the_node_bound():
    """A context whose IPM tile is the blocked one bounds cfg3-sized nodes
    with the same decisions and bounds within 1e-6 of the register tile's."""
    SP = O.build_L(I.gnp_adj(200, 0.1, 11, weights="int10"))
    P = api.make_params(include_Pent=0, include_Hepta=0, max_outer_iter=2, extra_iter=0)
    xf, sx = I.random_fixings(200, 30, 3, 12)
    seeds = np.arange(1, 4, dtype=np.uint64)
    outs = []
    for tile in (0, 22401):
        with forced(tile):
            ctx = api.BoundContext(SP, params=P, max_nodes=3)
        outs.append(ctx.bound(xf, sx, LB=0.0, seeds=seeds))
        ctx.close()
    a, b = outs
    assert np.all(a["status"] == 0) and np.all(b["status"] == 0)
    assert np.all(np.abs(a["bound"] - b["bound"]) <= 1e-6 * np.abs(a["bound"]))
    assert np.array_equal(a["stats"][:, :3], b["stats"][:, :3])


def test_autotuner_times_and_installs_a_valid_table():
    """The tile autotuner times both kernels for the classes it covers and
    installs a table of valid tiles; contexts then bound with whatever it
    chose (bounds within 1e-6 of the default table's)."""
    default = api.ipm_tile_table()
    try:
        table, timed = api.autotune_ipm(nmin=200, nmax=256, nodes=8, reps=2)
        print(table, timed, flush=True)
        assert set(timed) == {224, 256}
        reg = {224: 3207, 256: 3208}
        for ld, (mr, mb) in timed.items():
            assert mr > 0 and mb > 0
            assert table[ld] in (ld * 100 + 1, reg[ld])
        for ld in (64, 80, 96, 112, 128, 160, 192):
            assert table[ld] == default[ld]
        Ls = sub_matrices(250, 0.1, "int10", 31, 3, 40)
        compare(Ls, api.ipm_batch(Ls))
    finally:
        api.set_ipm_tile_table(default)


def test_failed_blocked_solves_are_repeated_with_the_register_tile():
    """cfg4-like QUBO nodes whose last iterations are ill-conditioned (the
    absolute gap test 1e-2 at |phi| ~ 4e4 asks for X to ~2e-7 relative):
    every solve ends with status 0 and phi within 1e-6 of the oracle, the
    blocked kernel or its register-tile repeat."""
    Q = I.random_qubo(150, 0.5, 7)
    Ls = sub_matrices_from_L(O.build_L(I.qubo_to_maxcut_adj(Q)), 8, 40, 8)
    with forced(16001):
        r = api.ipm_batch(Ls)
    for b, L in enumerate(Ls):
        _, _, phi, st, _ = O.ipm(L)
        assert r["status"][b] == 0
        assert abs(r["phi"][b] - phi) <= 1e-6 * abs(phi), (b, r["phi"][b], phi)
