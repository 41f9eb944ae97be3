"""Python mirror of the bounding entry points (calls the C ABI of
libbiqbin_gpu.so; see include/biqbin_gpu.h)."""
import ctypes as C

import numpy as np

from . import _native as N


def declare(L):
    L.bqb_ipm_debug.argtypes = [C.c_int, N.dptr, C.c_int, N.dptr, N.dptr, N.dptr, N.dptr, N.dptr,
                                N.dptr, N.iptr, N.dptr, N.dptr]
    L.bqb_ipm_debug.restype = C.c_int
    L.bqb_dist_selftest.argtypes = [C.c_int, C.c_int, C.c_int, N.dptr, N.dptr, N.iptr]
    L.bqb_dist_selftest.restype = C.c_int
    L.bqb_ipm_ld_for.argtypes = [C.c_int]
    L.bqb_ipm_ld_for.restype = C.c_int


def _d(a):
    return a.ctypes.data_as(N.dptr)


def _i(a):
    return a.ctypes.data_as(N.iptr)


def ipm_mc_pk(L):
    """Drop-in of src/ipm_mc_pk.c:28: returns (X, phi) for one matrix."""
    L = np.asfortranarray(L, dtype=np.float64)
    n = L.shape[0]
    X = np.zeros((n, n), order="F")
    phi = C.c_double()
    N.lib().ipm_mc_pk(_d(L), n, _d(X), C.byref(phi), 0)
    return X, phi.value


def ipm_batch(Ls, trace_cap=0):
    """Batched IPM over a list of square matrices.  Returns dict with X (list),
    phi, iters, trials, status, y (list) and alpha traces."""
    ns = np.array([m.shape[0] for m in Ls], dtype=np.int32)
    flat = np.concatenate([np.asfortranarray(m, dtype=np.float64).ravel(order="F") for m in Ls])
    tot = int(flat.size)
    X = np.zeros(tot)
    nodes = len(Ls)
    phi = np.zeros(nodes)
    iters = np.zeros(nodes, dtype=np.int32)
    trials = np.zeros(nodes, dtype=np.int32)
    status = np.zeros(nodes, dtype=np.int32)
    y = np.zeros(int(ns.sum()))
    tr = np.zeros(max(1, nodes * 2 * trace_cap))
    N.check(N.lib().bqb_ipm_batch(nodes, _i(ns), _d(flat), _d(X), _d(phi), _i(iters), _i(trials),
                                  _i(status), _d(y), _d(tr), trace_cap))
    Xs, ys, o, oy = [], [], 0, 0
    for n in ns:
        Xs.append(X[o:o + n * n].reshape((n, n), order="F"))
        ys.append(y[oy:oy + n])
        o += n * n
        oy += n
    traces = [tr[b * 2 * trace_cap: b * 2 * trace_cap + 2 * iters[b]] for b in range(nodes)] if trace_cap else None
    return dict(X=Xs, phi=phi, iters=iters, trials=trials, status=status, y=ys, trace=traces)


def ipm_debug(L, max_iters):
    """Internal state after `max_iters` IPM iterations (diagnostic entry)."""
    L = np.asfortranarray(L, dtype=np.float64)
    n = L.shape[0]
    X, Zi, Mi, dX = (np.zeros((n, n), order="F") for _ in range(4))
    y = np.zeros(n)
    phi = C.c_double()
    it = C.c_int()
    tr = np.zeros(2 * max_iters)
    dbg = np.zeros(4 * 256)
    N.check(N.lib().bqb_ipm_debug(n, _d(L), max_iters, _d(X), _d(Zi), _d(Mi), _d(dX), _d(y),
                                  C.byref(phi), C.byref(it), _d(tr), _d(dbg)))
    ld = N.lib().bqb_ipm_ld_for(n)
    return dict(X=X, Zi=Zi, Mi=Mi, dX=dX, y=y, phi=phi.value, iters=it.value, trace=tr,
                y0=dbg[:n], zd0=dbg[ld:ld + n], dgL=dbg[2 * ld:2 * ld + n], scal=dbg[3 * ld:3 * ld + 3])


def dist_selftest(tile, op, A):
    """Apply one distributed primitive (diagnostic entry); returns (out, ok)."""
    A = np.asfortranarray(A, dtype=np.float64)
    n = A.shape[0]
    out = np.zeros((n, n), order="F")
    ok = C.c_int()
    N.check(N.lib().bqb_dist_selftest(tile, op, n, _d(A), _d(out), C.byref(ok)))
    return out, ok.value
