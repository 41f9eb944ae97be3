"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where the oracle is the checker (or the timed
CPU baseline), never the product path.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

dptr = C.POINTER(C.c_double)
iptr = C.POINTER(C.c_int)


class Params(C.Structure):
    """Same fields/order as the reference's PARAM_FIELDS (src/biqbin.h:49-78)."""
    _fields_ = [
        ("init_bundle_iter", C.c_int), ("max_bundle_iter", C.c_int), ("triag_iter", C.c_int),
        ("pent_iter", C.c_int), ("hept_iter", C.c_int), ("max_outer_iter", C.c_int),
        ("extra_iter", C.c_int), ("violated_TriIneq", C.c_double), ("TriIneq", C.c_int),
        ("adjust_TriIneq", C.c_int), ("PentIneq", C.c_int), ("HeptaIneq", C.c_int),
        ("Pent_Trials", C.c_int), ("Hepta_Trials", C.c_int), ("include_Pent", C.c_int),
        ("include_Hepta", C.c_int), ("root", C.c_int), ("use_diff", C.c_int),
        ("time_limit", C.c_int), ("branchingStrategy", C.c_int),
    ]


class Tri(C.Structure):
    _fields_ = [("i", C.c_int), ("j", C.c_int), ("k", C.c_int), ("type", C.c_int),
                ("value", C.c_double), ("y", C.c_double)]


class Pent(C.Structure):
    _fields_ = [("type", C.c_int), ("perm", C.c_int * 5), ("value", C.c_double), ("y", C.c_double)]


class Hept(C.Structure):
    _fields_ = [("type", C.c_int), ("perm", C.c_int * 7), ("value", C.c_double), ("y", C.c_double)]


class IpmStats(C.Structure):
    _fields_ = [("iters", C.c_int), ("chol_trials", C.c_int), ("status", C.c_int)]


class NodeIn(C.Structure):
    _fields_ = [("xfixed", iptr), ("solX", iptr), ("LB", C.c_double), ("diff", C.c_double),
                ("is_root", C.c_int), ("seed", C.c_uint64)]


class NodeOut(C.Structure):
    _fields_ = [("bound", C.c_double), ("basic_bound", C.c_double), ("LB", C.c_double),
                ("improved", C.c_int), ("diff", C.c_double), ("prune", C.c_int),
                ("giveup", C.c_int), ("outer_iters", C.c_int), ("ipm_calls", C.c_int),
                ("ipm_iters", C.c_int), ("max_cuts", C.c_int), ("max_bundle", C.c_int),
                ("status", C.c_int)]


class BabOut(C.Structure):
    _fields_ = [("best", C.c_double), ("root_bound", C.c_double), ("nodes", C.c_int),
                ("status", C.c_int), ("seconds", C.c_double)]


def lib():
    """The oracle library in use: liboracle.so, or the variant selected by
    variant() (tests that show their checks detect a changed constant)."""
    global _LIB
    if _VARIANT is not None:
        return _VARIANT
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        _LIB = load(path)
    return _LIB


_VARIANT = None


class variant:
    """Context manager: every wrapper of this module calls the library at
    `path` (an oracle built with other constants) inside the block."""

    def __init__(self, path):
        self.L = load(path)

    def __enter__(self):
        global _VARIANT
        self.prev, _VARIANT = _VARIANT, self.L
        return self.L

    def __exit__(self, *exc):
        global _VARIANT
        _VARIANT = self.prev
        return False


def load(path):
    """Load an oracle shared library and declare its signatures."""
    L = C.CDLL(path)
    L.orc_ipm_mc_pk_ex.argtypes = [dptr, C.c_int, dptr, dptr, dptr, C.POINTER(IpmStats), dptr, C.c_int]
    L.orc_op_B.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, dptr, dptr]
    L.orc_op_Bt.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, dptr, dptr]
    L.orc_get_violated_triangles.argtypes = [dptr, C.c_int, C.c_double, C.c_int, C.POINTER(Tri), iptr]
    L.orc_get_violated_triangles.restype = C.c_double
    L.orc_qap_sa.argtypes = [iptr, C.c_int, dptr, C.c_int, iptr, C.c_uint64, iptr]
    L.orc_qap_sa.restype = C.c_double
    L.orc_H_pent.argtypes = [C.c_int]
    L.orc_H_pent.restype = iptr
    L.orc_H_hept.argtypes = [C.c_int]
    L.orc_H_hept.restype = iptr
    L.orc_get_violated_pent.argtypes = [dptr, C.c_int, C.c_int, C.c_int, C.POINTER(Pent), iptr, C.c_uint64, iptr]
    L.orc_get_violated_pent.restype = C.c_double
    L.orc_get_violated_hept.argtypes = [dptr, C.c_int, C.c_int, C.c_int, C.POINTER(Hept), iptr, C.c_uint64, iptr]
    L.orc_get_violated_hept.restype = C.c_double
    L.orc_mc_1opt.argtypes = [iptr, dptr, C.c_int]
    L.orc_mc_1opt.restype = C.c_double
    L.orc_evaluate_solution.argtypes = [dptr, C.c_int, iptr]
    L.orc_evaluate_solution.restype = C.c_double
    L.orc_create_subproblem.argtypes = [dptr, C.c_int, iptr, iptr, dptr]
    L.orc_create_subproblem.restype = C.c_int
    L.orc_get_fixed_value.argtypes = [dptr, C.c_int, iptr, iptr]
    L.orc_get_fixed_value.restype = C.c_double
    L.orc_build_L_from_adj.argtypes = [dptr, C.c_int, dptr]
    L.orc_solve_lambda.argtypes = [C.c_int, dptr, dptr, dptr]
    L.orc_solve_lambda.restype = C.c_int
    L.orc_lambda_eta.argtypes = [C.c_int, C.c_int, dptr, dptr, C.c_int, dptr, dptr, dptr, dptr, dptr]
    L.orc_lambda_eta.restype = C.c_int
    L.orc_fct_eval.argtypes = [dptr, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                               dptr, dptr, dptr, C.POINTER(IpmStats)]
    L.orc_fct_eval.restype = C.c_double
    L.orc_ws_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(Params)]
    L.orc_ws_create.restype = C.c_void_p
    L.orc_ws_destroy.argtypes = [C.c_void_p]
    L.orc_sdpbound.argtypes = [C.c_void_p, dptr, C.c_int, C.POINTER(NodeIn), C.POINTER(NodeOut), dptr, iptr]
    L.orc_bab_solve.argtypes = [dptr, C.c_int, C.POINTER(Params), C.c_uint64, C.c_int, iptr, C.POINTER(BabOut)]
    L.orc_params_default.argtypes = [C.POINTER(Params)]
    L.orc_sdpbound_many.argtypes = [dptr, C.c_int, C.POINTER(Params), C.c_int, iptr, iptr, dptr, dptr,
                                    C.POINTER(C.c_uint64), C.c_int, dptr, iptr]
    L.orc_key.argtypes = [C.c_uint64, C.c_uint64]
    L.orc_key.restype = C.c_uint64
    L.orc_draw.argtypes = [C.c_uint64, C.c_uint64]
    L.orc_draw.restype = C.c_uint64
    L.orc_heuristic_on.argtypes = [dptr, C.c_int, iptr, iptr, dptr, dptr, C.c_int, C.c_uint64, iptr, iptr]
    L.orc_heuristic_on.restype = C.c_double
    L.orc_ws_set_trace.argtypes = [C.c_void_p, dptr, C.c_int]
    L.orc_ws_trace_n.argtypes = [C.c_void_p]
    L.orc_ws_trace_n.restype = C.c_int
    L.orc_chol_upper.argtypes = [dptr, C.c_int]
    L.orc_chol_upper.restype = C.c_int
    L.orc_update_triangles_on.argtypes = [dptr, C.c_int, C.POINTER(Params), C.POINTER(Tri), iptr, C.c_int,
                                          dptr, C.POINTER(Tri), iptr]
    L.orc_update_triangles_on.restype = C.c_double
    return L


def _d(a):
    return a.ctypes.data_as(dptr)


def _i(a):
    return a.ctypes.data_as(iptr)


def f64(a):
    """Column-major contiguous float64 copy (Fortran order) of a square matrix."""
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def default_params(**kw):
    p = Params()
    lib().orc_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def reference_params_file(**kw):
    """The reference's shipped `params` file values (differ from the compiled defaults)."""
    d = dict(init_bundle_iter=5, max_bundle_iter=15, triag_iter=5, pent_iter=5, hept_iter=5,
             max_outer_iter=20, extra_iter=10, violated_TriIneq=0.05, TriIneq=5000, adjust_TriIneq=1,
             PentIneq=5000, HeptaIneq=5000, Pent_Trials=60, Hepta_Trials=50, include_Pent=1,
             include_Hepta=1, root=0, use_diff=1, time_limit=0, branchingStrategy=1)
    d.update(kw)
    return default_params(**d)


def ipm(L, trace_cap=0):
    """Returns (X, y, phi, stats, alpha_trace)."""
    L = f64(L)
    n = L.shape[0]
    X = np.zeros((n, n), order="F")
    y = np.zeros(n)
    phi = C.c_double()
    st = IpmStats()
    tr = np.zeros(max(trace_cap, 1))
    lib().orc_ipm_mc_pk_ex(_d(L), n, _d(X), _d(y), C.byref(phi), C.byref(st), _d(tr), trace_cap)
    return X, y, phi.value, st, tr[: 2 * st.iters] if trace_cap else None


def cut_arrays(tris=(), pents=(), hepts=()):
    """Build ctypes arrays of cut records from python tuples
    tri: (i, j, k, type); pent: (type, perm5); hept: (type, perm7)."""
    ta = (Tri * max(len(tris), 1))()
    for c, (i, j, k, t) in enumerate(tris):
        ta[c].i, ta[c].j, ta[c].k, ta[c].type = i, j, k, t
    pa = (Pent * max(len(pents), 1))()
    for c, (t, p) in enumerate(pents):
        pa[c].type = t
        for a in range(5):
            pa[c].perm[a] = p[a]
    ha = (Hept * max(len(hepts), 1))()
    for c, (t, p) in enumerate(hepts):
        ha[c].type = t
        for a in range(7):
            ha[c].perm[a] = p[a]
    return ta, len(tris), pa, len(pents), ha, len(hepts)


def op_B(X, cuts, y=None):
    X = f64(X)
    n = X.shape[0]
    ta, nt, pa, np_, ha, nh = cuts
    m = nt + np_ + nh
    y = np.zeros(m) if y is None else np.array(y, dtype=np.float64)
    lib().orc_op_B(n, ta, nt, pa, np_, ha, nh, _d(y), _d(X))
    return y


def op_Bt(X, cuts, t):
    X = f64(X).copy(order="F")
    n = X.shape[0]
    ta, nt, pa, np_, ha, nh = cuts
    t = np.ascontiguousarray(t, dtype=np.float64)
    lib().orc_op_Bt(n, ta, nt, pa, np_, ha, nh, _d(X), _d(t))
    return X


def triangles(X, threshold, cap):
    X = f64(X)
    n = X.shape[0]
    lst = (Tri * max(cap, 1))()
    size = C.c_int()
    mx = lib().orc_get_violated_triangles(_d(X), n, threshold, cap, lst, C.byref(size))
    return mx, [(lst[c].i, lst[c].j, lst[c].k, lst[c].type, lst[c].value) for c in range(size.value)]


def fct_eval(L, cuts, gamma):
    L = f64(L)
    n = L.shape[0]
    ta, nt, pa, np_, ha, nh = cuts
    m = nt + np_ + nh
    X = np.zeros((n, n), order="F")
    g = np.zeros(max(m, 1))
    gamma = np.ascontiguousarray(gamma, dtype=np.float64)
    st = IpmStats()
    f = lib().orc_fct_eval(_d(L), n, ta, nt, pa, np_, ha, nh, _d(gamma), _d(X), _d(g), C.byref(st))
    return f, X, g[:m], st


def build_L(adj):
    adj = f64(adj)
    n0 = adj.shape[0]
    L = np.zeros((n0, n0), order="F")
    lib().orc_build_L_from_adj(_d(adj), n0, _d(L))
    return L


def create_subproblem(SP_L, xfixed, solX):
    SP_L = f64(SP_L)
    n0 = SP_L.shape[0]
    xf = np.ascontiguousarray(xfixed, dtype=np.int32)
    sx = np.ascontiguousarray(solX, dtype=np.int32)
    PL = np.zeros(n0 * n0)
    n = lib().orc_create_subproblem(_d(SP_L), n0, _i(xf), _i(sx), _d(PL))
    fv = lib().orc_get_fixed_value(_d(SP_L), n0, _i(xf), _i(sx))
    return PL[: n * n].reshape((n, n), order="F"), fv


def evaluate_solution(SP_L, sol):
    SP_L = f64(SP_L)
    s = np.ascontiguousarray(sol, dtype=np.int32)
    return lib().orc_evaluate_solution(_d(SP_L), SP_L.shape[0], _i(s))


def sdpbound(SP_L, xfixed, solX, LB, diff, is_root, seed, params, ws=None):
    SP_L = f64(SP_L)
    n0 = SP_L.shape[0]
    own = ws is None
    if own:
        ws = lib().orc_ws_create(n0, 400, 50000, 50000, 50000, C.byref(params))
    xf = np.ascontiguousarray(xfixed, dtype=np.int32)
    sx = np.ascontiguousarray(solX, dtype=np.int32)
    nin = NodeIn(_i(xf), _i(sx), LB, diff, int(is_root), seed)
    out = NodeOut()
    frac = np.zeros(n0 - 1)
    xb = np.zeros(n0 - 1, dtype=np.int32)
    lib().orc_sdpbound(ws, _d(SP_L), n0, C.byref(nin), C.byref(out), _d(frac), _i(xb))
    if own:
        lib().orc_ws_destroy(ws)
    return out, frac, xb


def lambda_eta(zeta, G, gamma, t):
    """lambda_eta (bundle.c:292-373) for G (m x k, column-major), returns
    (dgamma, lambda, eta, t_out, status)."""
    G = f64(G)
    m, k = G.shape
    zeta = np.ascontiguousarray(zeta, dtype=np.float64)
    gamma = np.ascontiguousarray(gamma, dtype=np.float64)
    dg, eta, lam = np.zeros(max(m, 1)), np.zeros(max(m, 1)), np.zeros(k)
    tt = C.c_double(t)
    st = lib().orc_lambda_eta(m, k, _d(zeta), _d(G), m, _d(gamma), _d(dg), _d(lam), _d(eta), C.byref(tt))
    return dg[:m], lam, eta[:m], tt.value, st


def update_triangles_on(X, params, cuts, y, max_tri):
    """updateTriangleInequalities (cutting_planes.c:132-214) on X and a cut
    list of (i, j, k, type) with multipliers y.  Returns (cuts_out as
    (i, j, k, type, y) tuples, list as (i, j, k, type, value), max violation)."""
    X = f64(X)
    n = X.shape[0]
    ca = (Tri * max(max_tri, 1))()
    for c, (i, j, k, t) in enumerate(cuts):
        ca[c].i, ca[c].j, ca[c].k, ca[c].type = i, j, k, t
    nc = C.c_int(len(cuts))
    la = (Tri * max(params.TriIneq, 1))()
    ls = C.c_int()
    y = np.ascontiguousarray(np.asarray(y, dtype=np.float64).reshape(-1) if len(cuts) else np.zeros(1))
    mx = lib().orc_update_triangles_on(_d(X), n, C.byref(params), ca, C.byref(nc), max_tri, _d(y), la,
                                       C.byref(ls))
    out = [(ca[c].i, ca[c].j, ca[c].k, ca[c].type, ca[c].y) for c in range(nc.value)]
    lst = [(la[c].i, la[c].j, la[c].k, la[c].type, la[c].value) for c in range(ls.value)]
    return out, lst, mx


def sdpbound_trace(SP_L, xfixed, solX, LB, diff, is_root, seed, params, cap=40000):
    """sdpbound with the step trace: returns (out, frac, xbest, rows[r, 8])."""
    SP_L = f64(SP_L)
    ws = lib().orc_ws_create(SP_L.shape[0], 400, 50000, 50000, 50000, C.byref(params))
    buf = np.zeros(8 * cap)
    lib().orc_ws_set_trace(ws, _d(buf), cap)
    try:
        out, frac, xb = sdpbound(SP_L, xfixed, solX, LB, diff, is_root, seed, params, ws=ws)
        rows = buf[: 8 * lib().orc_ws_trace_n(ws)].reshape(-1, 8).copy()
    finally:
        lib().orc_ws_destroy(ws)
    return out, frac, xb, rows


def sdpbound_many(SP_L, params, xfixed, solX, LB, diff, seeds, nthreads):
    """Bound many nodes with nthreads CPU threads; returns (bounds, status)."""
    SP_L = f64(SP_L)
    n0 = SP_L.shape[0]
    xf = np.ascontiguousarray(xfixed, dtype=np.int32)
    sx = np.ascontiguousarray(solX, dtype=np.int32)
    count = xf.shape[0]
    lb = np.ascontiguousarray(np.broadcast_to(LB, (count,)), dtype=np.float64)
    df = np.ascontiguousarray(np.broadcast_to(diff, (count,)), dtype=np.float64)
    sd = np.ascontiguousarray(np.broadcast_to(seeds, (count,)), dtype=np.uint64)
    bounds = np.zeros(count)
    status = np.zeros(count, dtype=np.int32)
    lib().orc_sdpbound_many(_d(SP_L), n0, C.byref(params), count, _i(xf), _i(sx), _d(lb), _d(df),
                            sd.ctypes.data_as(C.POINTER(C.c_uint64)), nthreads, _d(bounds), _i(status))
    return bounds, status


def bab(SP_L, params, seed=2020, batch=1):
    SP_L = f64(SP_L)
    n0 = SP_L.shape[0]
    sol = np.zeros(n0 - 1, dtype=np.int32)
    out = BabOut()
    lib().orc_bab_solve(_d(SP_L), n0, C.byref(params), seed, batch, _i(sol), C.byref(out))
    return out, sol


def heuristic_on(SP_L, xfixed, solX, PL, X, key):
    SP_L = f64(SP_L)
    n0 = SP_L.shape[0]
    PL = f64(PL)
    X = f64(X)
    xf = np.ascontiguousarray(xfixed, dtype=np.int32)
    sx = np.ascontiguousarray(solX, dtype=np.int32)
    x = np.zeros(n0 - 1, dtype=np.int32)
    st = C.c_int()
    v = lib().orc_heuristic_on(_d(SP_L), n0, _i(xf), _i(sx), _d(PL), _d(X), PL.shape[0], key, _i(x), C.byref(st))
    return v, x, st.value


def key(a, b):
    return lib().orc_key(a, b)


def pent_sep(X, trials, cap, k):
    X = f64(X)
    n = X.shape[0]
    lst = (Pent * max(cap, 1))()
    size = C.c_int()
    st = C.c_int()
    mn = lib().orc_get_violated_pent(_d(X), n, trials, cap, lst, C.byref(size), k, C.byref(st))
    return mn, [(lst[c].type, [lst[c].perm[a] for a in range(5)], lst[c].value) for c in range(size.value)], st.value


def hept_sep(X, trials, cap, k):
    X = f64(X)
    n = X.shape[0]
    lst = (Hept * max(cap, 1))()
    size = C.c_int()
    st = C.c_int()
    mn = lib().orc_get_violated_hept(_d(X), n, trials, cap, lst, C.byref(size), k, C.byref(st))
    return mn, [(lst[c].type, [lst[c].perm[a] for a in range(7)], lst[c].value) for c in range(size.value)], st.value
