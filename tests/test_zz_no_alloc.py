"""bqb_ipm_batch_noalloc (the batched IPM that takes its work memory from the
library's static device array) against bqb_ipm_batch, through the C ABI.

Device buffers for inputs and outputs are allocated by the test through
the library's own HIP runtime.  The
two entry points launch the same kernel with the same register tile, so phi,
X, iteration counts and statuses must agree bit for bit.  Two calls in a
row, the second smaller, show that nothing of the first call leaks into the
second; a request larger than the scratch array is rejected."""
import ctypes as C

import numpy as np
import pytest

import oracle_py as O
from devmem import DevArray
from parallel_biqbin_maxcut_amd import _native as N, api, instances as I

pytestmark = pytest.mark.gpu


def _subproblems(n0, p, weights, seed, count, nfix_max):
    adj = I.gnp_adj(n0, p, seed, weights=weights)
    L = O.build_L(adj)
    xf, sx = I.random_fixings(n0, nfix_max, count, seed + 1)
    return [O.create_subproblem(L, xf[b], sx[b])[0] for b in range(count)]


def _noalloc(Ls):
    ns = np.array([m.shape[0] for m in Ls], dtype=np.int32)
    flat = np.concatenate([np.asfortranarray(m, dtype=np.float64).ravel(order="F") for m in Ls])
    nodes = len(Ls)
    print(f"noalloc call: {nodes} nodes, n {ns.min()}..{ns.max()}", flush=True)
    n_d, L_d = DevArray(ns), DevArray(flat)
    X_d = DevArray(like=flat, fill_byte=0xFF)            # NaN until written
    phi_d = DevArray(like=np.zeros(nodes))
    it_d = DevArray(like=np.zeros(nodes, np.int32))
    st_d = DevArray(like=np.zeros(nodes, np.int32), fill_byte=0xFF)
    bufs = (n_d, L_d, X_d, phi_d, it_d, st_d)
    try:
        N.check(N.lib().bqb_ipm_batch_noalloc(nodes, int(ns.max()), *(b.ptr for b in bufs), None))
        X = X_d.get()
        r = dict(phi=phi_d.get(), iters=it_d.get(), status=st_d.get())
    finally:
        for b in bufs:
            b.free()
    Xs, o = [], 0
    for n in ns:
        Xs.append(X[o:o + n * n].reshape((n, n), order="F"))
        o += n * n
    r["X"] = Xs
    return r


def _same(Ls, a, b):
    assert np.array_equal(a["status"], b["status"]) and not a["status"].any()
    assert np.array_equal(a["iters"], b["iters"])
    assert np.array_equal(a["phi"], b["phi"]), (a["phi"], b["phi"])
    for k in range(len(Ls)):
        assert np.array_equal(a["X"][k], b["X"][k]), k


def test_noalloc_matches_allocating_entry_twice():
    N.check(N.lib().bqb_init(0))
    big = _subproblems(100, 0.3, "pm1", 5, 24, 40)          # sizes 60-100
    small = _subproblems(40, 0.5, "int10", 9, 5, 10)        # sizes 30-40, fewer nodes
    for Ls in (big, small):
        ref = api.ipm_batch(Ls)
        got = _noalloc(Ls)
        _same(Ls, got, ref)


def test_noalloc_rejects_oversized_request():
    N.check(N.lib().bqb_init(0))
    lib = N.lib()
    need = lib.bqb_ipm_noalloc_bytes(64, 250)
    assert need > (64 << 20)
    rc = lib.bqb_ipm_batch_noalloc(64, 250, C.c_void_p(16), C.c_void_p(16), C.c_void_p(16), C.c_void_p(16),
                                   C.c_void_p(16), C.c_void_p(16), None)
    assert rc == -1
    assert b"scratch" in lib.bqb_last_error()
