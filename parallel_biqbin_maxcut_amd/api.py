"""Python mirror of the bounding entry points (calls the C ABI of
libbiqbin_gpu.so; see include/biqbin_gpu.h)."""
import ctypes as C
import os

import numpy as np

from . import _native as N


def declare(L):
    _declare_host(L)
    L.bqb_ipm_debug.argtypes = [C.c_int, N.dptr, C.c_int, N.dptr, N.dptr, N.dptr, N.dptr, N.dptr,
                                N.dptr, N.iptr, N.dptr, N.dptr]
    L.bqb_ipm_debug.restype = C.c_int
    L.bqb_dist_selftest.argtypes = [C.c_int, C.c_int, C.c_int, N.dptr, N.dptr, N.iptr]
    L.bqb_dist_selftest.restype = C.c_int
    L.bqb_copy_probe.argtypes = [C.c_size_t, C.c_int, N.dptr]
    L.bqb_set_ipm_tile.argtypes = [C.c_int]
    L.bqb_ipm_ld_for.argtypes = [C.c_int]
    L.bqb_ipm_ld_for.restype = C.c_int
    vp = C.c_void_p
    L.bqb_ipm_batch_mixed.argtypes = [C.c_int, N.iptr, N.dptr, N.dptr, N.dptr, N.dptr, N.iptr, N.iptr, N.iptr,
                                      N.dptr, C.c_int, N.dptr]
    L.bqb_ipm_batch_mixed.restype = C.c_int
    L.bqb_mx_selftest.argtypes = [C.c_int, C.c_int, N.dptr, N.dptr, N.dptr]
    L.bqb_ipm_noalloc_bytes.argtypes = [C.c_int, C.c_int]
    L.bqb_ipm_noalloc_bytes.restype = C.c_size_t
    L.bqb_ipm_batch_noalloc.argtypes = [C.c_int, C.c_int, vp, vp, vp, vp, vp, vp, vp]
    L.bqb_ipm_batch_noalloc.restype = C.c_int
    L.bqb_ipm_tile_table.argtypes = [N.iptr]
    L.bqb_set_ipm_tile_table.argtypes = [N.iptr]
    L.bqb_autotune_ipm.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, N.dptr, N.iptr]
    L.bqb_ipm_tile_for.argtypes = [C.c_int]
    _declare_ctx(L)
    from . import learn
    learn.declare(L)


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


TILE_CLASS_LD = (64, 80, 96, 112, 128, 160, 192, 224, 256)


def ipm_tile_table():
    """{slab ld of the size class: IPM tile} (register tile P*100+R, or the
    blocked solve ld*100+1)."""
    t = np.zeros(len(TILE_CLASS_LD), dtype=np.int32)
    N.lib().bqb_ipm_tile_table(_i(t))
    return dict(zip(TILE_CLASS_LD, (int(v) for v in t)))


def set_ipm_tile_table(table):
    t = np.array([int(table[ld]) for ld in TILE_CLASS_LD], dtype=np.int32)
    N.check(N.lib().bqb_set_ipm_tile_table(_i(t)))


def autotune_ipm(nmin=1, nmax=256, nodes=64, reps=3):
    """Time the register tile against the blocked solve for every size class
    in [nmin, nmax] and install the faster (bqb_autotune_ipm).  Returns
    (table, {ld: (register ms, blocked ms)})."""
    ms = np.zeros(2 * len(TILE_CLASS_LD))
    t = np.zeros(len(TILE_CLASS_LD), dtype=np.int32)
    N.check(N.lib().bqb_autotune_ipm(nmin, nmax, nodes, reps, _d(ms), _i(t)))
    timed = {ld: (round(float(ms[2 * c]), 4), round(float(ms[2 * c + 1]), 4))
             for c, ld in enumerate(TILE_CLASS_LD) if ms[2 * c] > 0}
    return dict(zip(TILE_CLASS_LD, (int(v) for v in t))), timed


def ipm_auto_table():
    """{slab ld of the size class: (IPM tile, slices)} of contexts with
    automatic precision (slices 0: FP64 products)."""
    t = np.zeros(len(TILE_CLASS_LD), dtype=np.int32)
    sl = np.zeros(len(TILE_CLASS_LD), dtype=np.int32)
    N.lib().bqb_ipm_auto_table(_i(t), _i(sl))
    return {ld: (int(t[c]), int(sl[c])) for c, ld in enumerate(TILE_CLASS_LD)}


def set_ipm_auto_table(table):
    t = np.array([int(table[ld][0]) for ld in TILE_CLASS_LD], dtype=np.int32)
    sl = np.array([int(table[ld][1]) for ld in TILE_CLASS_LD], dtype=np.int32)
    N.check(N.lib().bqb_set_ipm_auto_table(_i(t), _i(sl)))


def autotune_ipm_auto(nmin=1, nmax=256, nodes=64, reps=3):
    """Time register / blocked solves with FP64 and bf16 products for every
    size class in [nmin, nmax] and install the fastest in the automatic
    table (bqb_autotune_ipm_auto).  Returns (table, {ld: (register FP64,
    blocked FP64, register bf16, blocked bf16) ms})."""
    ms = np.zeros(4 * len(TILE_CLASS_LD))
    t = np.zeros(len(TILE_CLASS_LD), dtype=np.int32)
    sl = np.zeros(len(TILE_CLASS_LD), dtype=np.int32)
    N.check(N.lib().bqb_autotune_ipm_auto(nmin, nmax, nodes, reps, _d(ms), _i(t), _i(sl)))
    timed = {ld: tuple(round(float(ms[4 * c + v]), 4) for v in range(4))
             for c, ld in enumerate(TILE_CLASS_LD) if ms[4 * c] > 0}
    return {ld: (int(t[c]), int(sl[c])) for c, ld in enumerate(TILE_CLASS_LD)}, timed


def mx_product(A, B, slices=3):
    """C = A B through the sliced bf16 matrix-core product (diagnostic)."""
    A = np.asfortranarray(A, dtype=np.float64)
    B = np.asfortranarray(B, dtype=np.float64)
    n = A.shape[0]
    Cm = np.zeros((n, n), order="F")
    N.check(N.lib().bqb_mx_selftest(n, slices, _d(A), _d(B), _d(Cm)))
    return Cm


def ipm_batch_mixed(Ls, slices=4):
    """Batched IPM with the products on the bf16 matrix cores (slices > 0)
    or in FP64 (slices = 0), and certified bounds phi_cert.  Returns dict
    with X (list), phi, phi_cert, iters, trials, status, y (list), ms."""
    ns = np.array([m.shape[0] for m in Ls], dtype=np.int32)
    flat = np.concatenate([np.asfortranarray(m, dtype=np.float64).ravel(order="F") for m in Ls])
    nodes = len(Ls)
    X = np.zeros(flat.size)
    phi, cert = np.zeros(nodes), np.zeros(nodes)
    iters, trials, status = (np.zeros(nodes, dtype=np.int32) for _ in range(3))
    y = np.zeros(int(ns.sum()))
    ms = C.c_double()
    N.check(N.lib().bqb_ipm_batch_mixed(nodes, _i(ns), _d(flat), _d(X), _d(phi), _d(cert), _i(iters), _i(trials),
                                        _i(status), _d(y), slices, C.byref(ms)))
    Xs, ys, o, oy = [], [], 0, 0
    for n in ns:
        Xs.append(X[o:o + n * n].reshape((n, n), order="F"))
        ys.append(y[oy:oy + n])
        o += n * n
        oy += n
    return dict(X=Xs, phi=phi, phi_cert=cert, iters=iters, trials=trials, status=status, y=ys, ms=ms.value)


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


def copy_probe(nbytes=256 << 20, reps=10):
    """Average ms of a plain device copy of nbytes (HBM counter calibration:
    each copy moves 2 x nbytes)."""
    ms = C.c_double()
    N.check(N.lib().bqb_copy_probe(nbytes, reps, C.byref(ms)))
    return ms.value



# ---------------------------------------------------------------- batch bound
class BiqBinParameters(C.Structure):
    """The reference's PARAM_FIELDS layout (src/biqbin.h:49-78)."""
    _fields_ = [
        ("init_bundle_iter", C.c_int), ("max_bundle_iter", C.c_int), ("triag_iter", C.c_int),
        ("pent_iter", C.c_int), ("hept_iter", C.c_int), ("max_outer_iter", C.c_int),
        ("extra_iter", C.c_int), ("violated_TriIneq", C.c_double), ("TriIneq", C.c_int),
        ("adjust_TriIneq", C.c_int), ("PentIneq", C.c_int), ("HeptaIneq", C.c_int),
        ("Pent_Trials", C.c_int), ("Hepta_Trials", C.c_int), ("include_Pent", C.c_int),
        ("include_Hepta", C.c_int), ("root", C.c_int), ("use_diff", C.c_int),
        ("time_limit", C.c_int), ("branchingStrategy", C.c_int),
    ]


DEFAULT_PARAMS = dict(init_bundle_iter=3, max_bundle_iter=15, triag_iter=5, pent_iter=5, hept_iter=5,
                      max_outer_iter=20, extra_iter=10, violated_TriIneq=1e-3, TriIneq=5000,
                      adjust_TriIneq=1, PentIneq=5000, HeptaIneq=5000, Pent_Trials=60,
                      Hepta_Trials=50, include_Pent=1, include_Hepta=1, root=0, use_diff=1,
                      time_limit=0, branchingStrategy=1)
# values of the reference's shipped `params` file
PARAMS_FILE = dict(DEFAULT_PARAMS, init_bundle_iter=5, violated_TriIneq=0.05)


def make_params(**kw):
    d = dict(PARAMS_FILE)
    d.update(kw)
    p = BiqBinParameters()
    for k, v in d.items():
        setattr(p, k, v)
    return p


class NodeState(C.Structure):
    _fields_ = [(f, C.c_int) for f in (
        "n", "status", "active", "prune", "giveup", "done", "count", "bdl_iter", "NIneq",
        "NPentIneq", "NHeptaIneq", "bundle", "sep_tri", "sep_pent", "sep_hept", "heur_idx",
        "improved", "ipm_calls", "ipm_iters", "max_cuts", "max_bundle_seen", "is_root", "serious",
        "cuts_changed", "off_pent", "off_hept", "stage", "bdl_left", "rphase", "flags", "warm", "cert_stale",
        "nsurv_tri", "nsurv_pent", "nsurv_hept")] + [(f, C.c_double) for f in (
        "fixedvalue", "f", "f_test", "t", "bound", "basic_bound", "oldf", "LB", "diff", "viol3",
        "viol5", "gap", "del_", "lmax", "f_cert")] + [("seed", C.c_uint64)]


BUF = dict(state=0, L=1, L0=2, X=3, Xt=4, tri=5, tri_y=6, pent=7, pent_y=8, hept=9, hept_y=10,
           gamma=11, gtest=12, g=13, G=14, F=15, lam=16, tri_list=17, tri_lval=18, kg_list=19,
           kg_lval=20, fracsol=21, xbest=22, xheur=23, mask=24, phi=25, Xb=26, nvec=27, dgamma=28,
           sa_val=29, ycen=30)
PH = dict(CERT_BASIC=110, CERT_CENTER=111, SETUP=0, POST_BASIC=1, CHECK_BASIC=2, BUNDLE_INIT=3, OUTER_BEGIN=4, OUTER_A=5, OUTER_B=6,
          SEP_PENT=7, SEP_HEPT=8, OUTER_C=9, REBUILD=10, FINISH=11, FCT_FINISH=12, LOAD=13,
          BUNDLE_NEXT=14, WARM=15, CERT_MASK=16, IPM_BASIC=100, IPM_L0=101, BASIC_TRAIN=115,
          HEUR=102, TRISEP=103, PENT=104, HEPT=105, CSR=106, OPBT=107, QP=108, STEP=109)


class Trajectory(C.Structure):
    """bqb_trajectory (include/biqbin_gpu.h)."""
    _fields_ = [("tag", C.c_longlong), ("n", C.c_int), ("rounds", C.c_int), ("ntri", C.c_int),
                ("npent", C.c_int), ("nhept", C.c_int), ("status", C.c_int), ("f", C.c_double),
                ("t", C.c_double), ("bound", C.c_double), ("fixedvalue", C.c_double),
                ("round_f", C.POINTER(C.c_double)), ("round_t", C.POINTER(C.c_double)),
                ("round_bound", C.POINTER(C.c_double)), ("round_cuts", C.POINTER(C.c_int)),
                ("tri", C.POINTER(C.c_int)), ("pent", C.POINTER(C.c_int)), ("hept", C.POINTER(C.c_int)),
                ("gamma", C.POINTER(C.c_double))]


def _arr(ptr, count, dtype, shape=None):
    if count == 0:
        a = np.zeros(0, dtype=dtype)
    else:
        a = np.ctypeslib.as_array(ptr, shape=(count,)).astype(dtype, copy=True)
    return a.reshape(shape) if shape is not None else a


class SolveOptions(C.Structure):
    """bqb_solve_options"""
    _fields_ = [("batch", C.c_int), ("max_bundle", C.c_int), ("slices", C.c_int), ("certify", C.c_int),
                ("model", C.c_void_p), ("drop", C.c_float), ("warm", C.c_int), ("graphs", C.c_int),
                ("ipm_warm", C.c_int)]


class SolveStats(C.Structure):
    _fields_ = [("best", C.c_double), ("root_bound", C.c_double), ("nodes", C.c_longlong),
                ("batches", C.c_longlong), ("status", C.c_int), ("stopped", C.c_int),
                ("max_queue", C.c_int), ("seconds", C.c_double), ("bound_ms", C.c_double),
                ("bundle_full", C.c_longlong), ("cuts_truncated", C.c_longlong)]


NMAX = 1024
NSTATS = 9   # ints per node in the stats arrays (BQB_NSTATS)
FLAG_BUNDLE_FULL, FLAG_TRI_TRUNC, FLAG_PENT_TRUNC, FLAG_HEPT_TRUNC, FLAG_FP64_RETRY, FLAG_REG_RETRY = 1, 2, 4, 8, 16, 32
FLAG_WARM_FAILED = 64


class Problem(C.Structure):
    """Layout of the reference's Problem (src/biqbin_cpp_api.h:27-35)."""
    _fields_ = [("L", C.POINTER(C.c_double)), ("n", C.c_int), ("NIneq", C.c_int),
                ("NPentIneq", C.c_int), ("NHeptaIneq", C.c_int), ("bundle", C.c_int)]


class BabNode(C.Structure):
    """Layout of the reference's BabNode (src/biqbin_cpp_api.h:54-63)."""
    _fields_ = [("xfixed", C.c_int * NMAX), ("solX", C.c_int * NMAX), ("fracsol", C.c_double * NMAX),
                ("level", C.c_int), ("upper_bound", C.c_double)]


def _declare_host(L):
    L.bqb_params_default.argtypes = [C.POINTER(BiqBinParameters)]
    L.bqb_params_default.restype = None
    L.bqb_read_params.argtypes = [C.c_char_p, C.POINTER(BiqBinParameters)]
    L.bqb_read_rudy.argtypes = [C.c_char_p, N.iptr]
    L.bqb_read_rudy.restype = C.c_void_p
    L.bqb_build_L.argtypes = [C.c_int, N.dptr, N.dptr]
    L.bqb_evaluate_solution.argtypes = [C.c_int, N.dptr, N.iptr]
    L.bqb_evaluate_solution.restype = C.c_double
    L.bqb_solve.argtypes = [C.c_int, N.dptr, C.POINTER(BiqBinParameters), C.c_uint64, C.c_int, N.iptr,
                            C.POINTER(SolveStats)]
    L.bqb_solve_cap.argtypes = [C.c_int, N.dptr, C.POINTER(BiqBinParameters), C.c_uint64, C.c_int, C.c_int,
                                N.iptr, C.POINTER(SolveStats)]
    L.bqb_solve_ex.argtypes = [C.c_int, N.dptr, C.POINTER(BiqBinParameters), C.c_uint64, C.c_int, C.c_int,
                               C.c_int, C.c_int, C.c_void_p, C.c_float, N.iptr, C.POINTER(SolveStats)]
    L.bqb_solve_opt.argtypes = [C.c_int, N.dptr, C.POINTER(BiqBinParameters), C.c_uint64,
                                C.POINTER(SolveOptions), N.iptr, C.POINTER(SolveStats)]
    for f in ("Evaluate", "SDPbound"):
        getattr(L, f).argtypes = [C.POINTER(BabNode), C.POINTER(Problem), C.POINTER(Problem), C.c_int]
        getattr(L, f).restype = C.c_double
    L.createSubproblem.argtypes = [C.POINTER(BabNode), C.POINTER(Problem), C.POINTER(Problem)]
    L.createSubproblem.restype = None
    L.getFixedValue.argtypes = [C.POINTER(BabNode), C.POINTER(Problem)]
    L.getFixedValue.restype = C.c_double
    L.fct_eval.argtypes = [C.POINTER(Problem), N.dptr, N.dptr, N.dptr]
    L.fct_eval.restype = C.c_double
    L.bqb_dropin_incumbent.argtypes = [N.iptr]
    L.bqb_dropin_incumbent.restype = C.c_double
    L.bqb_dropin_release.argtypes = []
    L.bqb_dropin_release.restype = None


def default_params():
    """The reference's PARAM_FIELDS defaults (src/biqbin.h:49-78)."""
    p = BiqBinParameters()
    N.lib().bqb_params_default(C.byref(p))
    return p


def read_params(path):
    """Parameter file as the reference reads it (src/process_input.c:153-180)."""
    p = BiqBinParameters()
    N.check(N.lib().bqb_read_params(os.fsencode(path), C.byref(p)))
    return p


def read_rudy(path):
    """Adjacency matrix of a rudy graph file (src/process_input.c:206-250)."""
    nv = C.c_int()
    ptr = N.lib().bqb_read_rudy(os.fsencode(path), C.byref(nv))
    if not ptr:
        raise ValueError(N.lib().bqb_last_error().decode())
    try:
        a = np.ctypeslib.as_array(C.cast(ptr, N.dptr), shape=(nv.value * nv.value,)).copy()
    finally:
        C.CDLL(None).free(C.c_void_p(ptr))
    return a.reshape((nv.value, nv.value), order="F")


def maxcut_L(adj):
    """SP->L of a graph (src/process_input.c:259-344), column-major."""
    adj = np.asfortranarray(adj, dtype=np.float64)
    nv = adj.shape[0]
    L = np.zeros((nv, nv), order="F")
    N.check(N.lib().bqb_build_L(nv, _d(adj), _d(L)))
    return L


def evaluate_solution(SP_L, sol):
    SP_L = np.asfortranarray(SP_L, dtype=np.float64)
    s = np.ascontiguousarray(sol, dtype=np.int32)
    return N.lib().bqb_evaluate_solution(SP_L.shape[0], _d(SP_L), _i(s))


def solve(SP_L, params=None, seed=2020, batch=64, max_bundle=0, slices=0, certify=0, model=None, drop=0.0,
          graphs=None, warm=False, ipm_warm=False):
    """Branch and bound on the GPU (bqb_solve_opt).  Returns (stats,
    solution) with the solution over the first n0-1 vertices (the last is on
    side 0).  graphs: launch trains as HIP graphs (None: the library's
    default).  warm: children inherit their parent's final cuts and
    multipliers.  ipm_warm: every bundle-step IPM starts from the node's
    previous one."""
    SP_L = np.asfortranarray(SP_L, dtype=np.float64)
    n0 = SP_L.shape[0]
    p = params if params is not None else make_params()
    sol = np.zeros(n0 - 1, dtype=np.int32)
    st = SolveStats()
    o = SolveOptions(batch, max_bundle, slices, certify, model.h if model is not None else None, drop,
                     1 if warm else 0, -1 if graphs is None else (1 if graphs else 0), 1 if ipm_warm else 0)
    N.check(N.lib().bqb_solve_opt(n0, _d(SP_L), C.byref(p), seed, C.byref(o), _i(sol), C.byref(st)))
    return st, sol


def _declare_ctx(L):
    vp = C.c_void_p
    L.bqb_ctx_create.argtypes = [C.c_int, N.dptr, C.POINTER(BiqBinParameters), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
    L.bqb_ctx_create.restype = vp
    L.bqb_ctx_destroy.argtypes = [vp]
    L.bqb_ctx_bytes.argtypes = [vp]
    L.bqb_ctx_bytes.restype = C.c_size_t
    L.bqb_ctx_ld.argtypes = [vp]
    L.bqb_ctx_ld.restype = C.c_int
    L.bqb_ctx_last_ms.argtypes = [vp]
    L.bqb_ctx_last_ms.restype = C.c_double
    L.bqb_abi_sizes.argtypes = [N.iptr, C.c_int]
    L.bqb_abi_sizes.restype = C.c_int
    u64p = C.POINTER(C.c_uint64)
    L.bqb_bound_batch.argtypes = [vp, C.c_int, N.iptr, N.iptr, N.dptr, N.dptr, N.iptr, u64p, N.dptr, N.dptr,
                                  N.iptr, N.dptr, N.iptr, N.dptr, N.iptr, N.iptr]
    L.bqb_fct_eval.argtypes = [vp, C.c_int, N.dptr, C.c_int, N.iptr, C.c_int, N.iptr, C.c_int, N.iptr, N.dptr,
                               N.dptr, N.dptr, N.dptr, N.iptr]
    L.bqb_bound_stream.argtypes = L.bqb_bound_batch.argtypes
    L.bqb_bound_stream_multi.argtypes = [C.POINTER(vp), C.c_int] + L.bqb_bound_batch.argtypes[1:]
    L.bqb_ctx_capture.argtypes = [vp, C.c_int]
    L.bqb_ctx_trajectory_count.argtypes = [vp]
    L.bqb_ctx_trajectory.argtypes = [vp, C.c_int, C.POINTER(Trajectory)]
    L.bqb_ctx_trajectory_clear.argtypes = [vp]
    L.bqb_ctx_trajectory_clear.restype = None
    L.bqb_ctx_push.argtypes = [vp, C.c_longlong, N.iptr, N.iptr, C.c_double, C.c_double, C.c_int, C.c_uint64]
    L.bqb_ctx_advance.argtypes = [vp]
    L.bqb_ctx_pop.argtypes = [vp, C.POINTER(C.c_longlong), N.dptr, N.dptr, N.iptr, N.dptr, N.iptr, N.dptr,
                              N.iptr, N.iptr]
    L.bqb_ctx_busy.argtypes = [vp]
    L.bqb_ctx_capacity.argtypes = [vp]
    L.bqb_ctx_set_warm.argtypes = [vp, C.c_int]
    L.bqb_ctx_set_ipm_warm.argtypes = [vp, C.c_int]
    L.bqb_ctx_push_cuts.argtypes = [vp, C.c_int, C.c_int, N.iptr, C.c_int, N.iptr, C.c_int, N.iptr, N.dptr]
    ipp = C.POINTER(N.iptr)
    L.bqb_ctx_last_cuts.argtypes = [vp, N.iptr, ipp, N.iptr, ipp, N.iptr, ipp, C.POINTER(N.dptr)]
    L.bqb_ctx_load.argtypes = [vp, C.c_int, N.iptr, N.iptr, N.dptr, N.dptr, N.iptr, u64p]
    L.bqb_ctx_run.argtypes = [vp, C.c_int, C.c_int]
    L.bqb_ctx_bound.argtypes = [vp]
    L.bqb_ctx_results.argtypes = [vp, N.dptr, N.dptr, N.iptr, N.dptr, N.iptr, N.dptr, N.iptr, N.iptr]
    L.bqb_ctx_download.argtypes = [vp, C.c_int, C.c_int, vp, C.c_size_t]
    L.bqb_ctx_upload.argtypes = [vp, C.c_int, C.c_int, vp, C.c_size_t]
    L.bqb_ctx_sync.argtypes = [vp]
    L.bqb_ctx_profile_ipm.argtypes = [vp, C.c_int]
    L.bqb_ctx_ipm_stats.argtypes = [vp, N.iptr, N.dptr, N.dptr]
    L.bqb_ctx_set_precision.argtypes = [vp, C.c_int, C.c_int]
    L.bqb_ctx_ipm_flops_mx.argtypes = [vp, N.dptr]
    L.bqb_ctx_set_graphs.argtypes = [vp, C.c_int]
    L.bqb_set_default_graphs.argtypes = [C.c_int]
    L.bqb_ctx_counters.argtypes = [vp, C.POINTER(C.c_longlong), C.c_int]


NCOUNTERS = 7
COUNTER_NAMES = ("advances", "round_slices", "eager_phases", "graph_launches", "graph_phases", "host_syncs",
                 "graph_instantiations")


def set_default_graphs(on):
    """Launch trains as HIP graphs in contexts created afterwards (bqb_solve
    included); returns the previous setting."""
    return N.lib().bqb_set_default_graphs(1 if on else 0)


class BoundContext:
    """Device context for bounding batches of B&B nodes of one instance."""

    def __init__(self, SP_L, params=None, max_nodes=1, max_bundle=0, max_tri=0, max_pent=0, max_hept=0,
                 groups=1):
        """groups > 1 adds slot groups of max_nodes slots each (contexts on
        their own streams); streamed bounding (bound(stream=True)) then runs
        through all groups with bqb_bound_stream_multi.  Everything else
        acts on the first group."""
        L = N.lib()
        self.SP_L = np.asfortranarray(SP_L, dtype=np.float64)
        self.n0 = self.SP_L.shape[0]
        self.params = params if params is not None else make_params()
        self.handles = []
        for _ in range(max(1, int(groups))):
            h = L.bqb_ctx_create(self.n0, _d(self.SP_L), C.byref(self.params), max_nodes, max_bundle,
                                 max_tri, max_pent, max_hept)
            if not h:
                err = L.bqb_last_error().decode()
                self.close()
                raise RuntimeError("bqb_ctx_create: " + err)
            self.handles.append(h)
        self.h = self.handles[0]
        self.ld = L.bqb_ctx_ld(self.h)
        self.max_nodes = max_nodes
        self.groups = len(self.handles)

    def close(self):
        for h in getattr(self, "handles", []):
            N.lib().bqb_ctx_destroy(h)
        self.handles = []
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _arrs(nodes, n0, xfixed, solX, LB, diff, is_root, seeds):
        B = n0 - 1
        xf = np.ascontiguousarray(np.reshape(xfixed, (nodes, B)), dtype=np.int32)
        sx = np.ascontiguousarray(np.reshape(solX, (nodes, B)), dtype=np.int32)
        lb = np.ascontiguousarray(np.broadcast_to(LB, (nodes,)), dtype=np.float64)
        df = np.ascontiguousarray(np.broadcast_to(diff, (nodes,)), dtype=np.float64)
        rt = np.ascontiguousarray(np.broadcast_to(is_root, (nodes,)), dtype=np.int32)
        sd = np.ascontiguousarray(np.broadcast_to(seeds, (nodes,)), dtype=np.uint64)
        return xf, sx, lb, df, rt, sd

    def bound(self, xfixed, solX, LB=0.0, diff=0.0, is_root=0, seeds=0, stream=False):
        """Bound the nodes (rows of xfixed/solX): bqb_bound_batch, or
        bqb_bound_stream (any count, slots refilled) if stream."""
        nodes = int(np.shape(xfixed)[0])
        xf, sx, lb, df, rt, sd = self._arrs(nodes, self.n0, xfixed, solX, LB, diff, is_root, seeds)
        B = self.n0 - 1
        out = dict(bound=np.zeros(nodes), fracsol=np.zeros((nodes, B)), xbest=np.zeros((nodes, B), np.int32),
                   LB=np.zeros(nodes), improved=np.zeros(nodes, np.int32), diff=np.zeros(nodes),
                   status=np.zeros(nodes, np.int32), stats=np.zeros((nodes, NSTATS), np.int32))
        args = (nodes, _i(xf), _i(sx), _d(lb), _d(df), _i(rt), sd.ctypes.data_as(C.POINTER(C.c_uint64)),
                _d(out["bound"]), _d(out["fracsol"]), _i(out["xbest"]), _d(out["LB"]), _i(out["improved"]),
                _d(out["diff"]), _i(out["status"]), _i(out["stats"]))
        if stream and self.groups > 1:
            hs = (C.c_void_p * self.groups)(*self.handles)
            N.check(N.lib().bqb_bound_stream_multi(hs, self.groups, *args))
        else:
            fn = N.lib().bqb_bound_stream if stream else N.lib().bqb_bound_batch
            N.check(fn(self.h, *args))
        out["ms"] = N.lib().bqb_ctx_last_ms(self.h)
        return out

    def load(self, xfixed, solX, LB=0.0, diff=0.0, is_root=0, seeds=0):
        nodes = int(np.shape(xfixed)[0])
        xf, sx, lb, df, rt, sd = self._arrs(nodes, self.n0, xfixed, solX, LB, diff, is_root, seeds)
        N.check(N.lib().bqb_ctx_load(self.h, nodes, _i(xf), _i(sx), _d(lb), _d(df), _i(rt),
                                     sd.ctypes.data_as(C.POINTER(C.c_uint64))))

    def run(self, phase, arg=0):
        N.check(N.lib().bqb_ctx_run(self.h, PH[phase] if isinstance(phase, str) else phase, arg))
        N.check(N.lib().bqb_ctx_sync(self.h))

    def get(self, buf, node, dtype, count):
        a = np.zeros(count, dtype=dtype)
        N.check(N.lib().bqb_ctx_download(self.h, BUF[buf], node, a.ctypes.data, a.nbytes))
        return a

    def put(self, buf, node, arr):
        a = np.ascontiguousarray(arr)
        N.check(N.lib().bqb_ctx_upload(self.h, BUF[buf], node, a.ctypes.data, a.nbytes))

    def state(self, node):
        s = NodeState()
        N.check(N.lib().bqb_ctx_download(self.h, 0, node, C.addressof(s), C.sizeof(s)))
        return s

    def set_state(self, node, s):
        N.check(N.lib().bqb_ctx_upload(self.h, 0, node, C.addressof(s), C.sizeof(s)))

    def fct_eval(self, L, tris=(), pents=(), hepts=(), gamma=()):
        """fct_eval (src/bundle.c:242-275) of matrix L with the given cuts
        (tri: (i, j, k, type); pent/hept: (type, vertices)).  Returns
        (f, X, g, status)."""
        L = np.asfortranarray(L, dtype=np.float64)
        n = L.shape[0]
        t = np.array([list(c) for c in tris] or [[0, 0, 0, 0]], dtype=np.int32)
        p = np.array([[c[0]] + list(c[1]) for c in pents] or [[0] * 6], dtype=np.int32)
        h = np.array([[c[0]] + list(c[1]) for c in hepts] or [[0] * 8], dtype=np.int32)
        m = len(tris) + len(pents) + len(hepts)
        gam = np.ascontiguousarray(np.asarray(gamma, dtype=np.float64).reshape(-1)[:m] if m else np.zeros(1))
        X = np.zeros((n, n), order="F")
        g = np.zeros(max(m, 1))
        f = C.c_double()
        st = C.c_int()
        N.check(N.lib().bqb_fct_eval(self.h, n, _d(L), len(tris), _i(t), len(pents), _i(p), len(hepts), _i(h),
                                     _d(gam), _d(X), _d(g), C.byref(f), C.byref(st)))
        return f.value, X, g[:m], st.value

    def sync(self):
        N.check(N.lib().bqb_ctx_sync(self.h))

    def capture(self, on=True):
        """Record the trajectory of every node that finishes (all groups)."""
        for h in self.handles:
            N.check(N.lib().bqb_ctx_capture(h, 1 if on else 0))

    def trajectories(self, clear=True):
        """The captured records as dicts of numpy arrays, ordered by tag."""
        L = N.lib()
        out = []
        for h in self.handles:
            for i in range(L.bqb_ctx_trajectory_count(h)):
                t = Trajectory()
                N.check(L.bqb_ctx_trajectory(h, i, C.byref(t)))
                R, m = t.rounds, t.ntri + t.npent + t.nhept
                out.append(dict(
                    tag=t.tag, n=t.n, rounds=R, status=t.status, f=t.f, t=t.t, bound=t.bound,
                    fixedvalue=t.fixedvalue, round_f=_arr(t.round_f, R, np.float64),
                    round_t=_arr(t.round_t, R, np.float64), round_bound=_arr(t.round_bound, R, np.float64),
                    round_cuts=_arr(t.round_cuts, 3 * R, np.int32, (R, 3)),
                    tri=_arr(t.tri, 4 * t.ntri, np.int32, (t.ntri, 4)),
                    pent=_arr(t.pent, 8 * t.npent, np.int32, (t.npent, 8)),
                    hept=_arr(t.hept, 8 * t.nhept, np.int32, (t.nhept, 8)),
                    gamma=_arr(t.gamma, m, np.float64)))
            if clear:
                L.bqb_ctx_trajectory_clear(h)
        out.sort(key=lambda r: r["tag"])
        return out

    def profile_ipm(self, max_launches):
        """Time every IPM launch (event pairs) and count its algorithmic
        flops on the device, for up to max_launches launches; 0 stops."""
        for h in self.handles:
            N.check(N.lib().bqb_ctx_profile_ipm(h, max_launches))

    def ipm_stats(self):
        """(launches, summed ms, algorithmic flops) since profiling started,
        over all slot groups."""
        tot = [0, 0.0, 0.0]
        for h in self.handles:
            k = C.c_int()
            ms = C.c_double()
            fl = C.c_double()
            N.check(N.lib().bqb_ctx_ipm_stats(h, C.byref(k), C.byref(ms), C.byref(fl)))
            tot[0] += k.value
            tot[1] += ms.value
            tot[2] += fl.value
        return tuple(tot)

    def ipm_flops_mx(self):
        """bf16 matrix-core flops since profiling started, over all groups."""
        tot = 0.0
        for h in self.handles:
            f = C.c_double()
            N.check(N.lib().bqb_ctx_ipm_flops_mx(h, C.byref(f)))
            tot += f.value
        return tot

    def set_graphs(self, on=True):
        for h in self.handles:
            N.check(N.lib().bqb_ctx_set_graphs(h, 1 if on else 0))

    def counters(self, reset=False):
        """Host-side launch and synchronisation counters summed over the slot
        groups (dict keyed by COUNTER_NAMES)."""
        tot = np.zeros(NCOUNTERS, dtype=np.int64)
        buf = (C.c_longlong * NCOUNTERS)()
        for h in self.handles:
            N.check(N.lib().bqb_ctx_counters(h, buf, 1 if reset else 0))
            tot += np.array(buf[:], dtype=np.int64)
        return dict(zip(COUNTER_NAMES, (int(v) for v in tot)))

    def set_precision(self, slices=0, certify=0):
        """IPM products in FP64 (slices 0) or on the bf16 matrix cores
        (2..4 slices), and certified bounds (implied by slices > 0)."""
        for h in self.handles:
            N.check(N.lib().bqb_ctx_set_precision(h, int(slices), int(certify)))

    def matrix(self, buf, node, n):
        a = self.get(buf, node, np.float64, self.ld * self.ld).reshape((self.ld, self.ld), order="F")
        return np.array(a[:n, :n], order="F")

    # ---- incremental use: push / advance / pop, inherited cuts ----------
    def set_warm(self, on=True):
        for h in self.handles:
            N.check(N.lib().bqb_ctx_set_warm(h, 1 if on else 0))

    def set_ipm_warm(self, on=True):
        """bqb_ctx_set_ipm_warm: bundle-step IPMs start from the slot's last."""
        for h in self.handles:
            N.check(N.lib().bqb_ctx_set_ipm_warm(h, 1 if on else 0))

    def push(self, xfixed, solX, LB=0.0, diff=0.0, is_root=0, seed=0, tag=0):
        """bqb_ctx_push of one node; returns its slot."""
        xf = np.ascontiguousarray(xfixed, dtype=np.int32)
        sx = np.ascontiguousarray(solX, dtype=np.int32)
        slot = N.lib().bqb_ctx_push(self.h, int(tag), _i(xf), _i(sx), float(LB), float(diff), int(is_root),
                                    int(seed))
        if slot < 0:
            raise RuntimeError("bqb_ctx_push: " + N.lib().bqb_last_error().decode())
        return slot

    def push_cuts(self, slot, tris=(), pents=(), hepts=(), gamma=()):
        """Inherited cuts of the node just pushed to `slot` (records as
        fct_eval's: tri (i, j, k, type), pent (type, 5 vertices), hept (type,
        7 vertices), vertices of the node's subproblem)."""
        t = np.ascontiguousarray(np.reshape(np.asarray(tris, dtype=np.int32), (-1, 4)))
        pn = np.zeros((len(pents), 8), np.int32)
        for q, r in enumerate(pents):
            pn[q, :6] = r
        hp = np.zeros((len(hepts), 8), np.int32)
        for q, r in enumerate(hepts):
            hp[q, :8] = r
        g = np.ascontiguousarray(gamma, dtype=np.float64)
        N.check(N.lib().bqb_ctx_push_cuts(self.h, int(slot), len(t), _i(t), len(pn), _i(pn), len(hp), _i(hp),
                                          _d(g)))

    def advance(self):
        N.check(N.lib().bqb_ctx_advance(self.h))

    def busy(self):
        return N.lib().bqb_ctx_busy(self.h)

    def pop(self):
        """One finished node (dict) or None; 'cuts' holds its final cut set
        (tri [k,4], pent [k,8], hept [k,8], gamma) in warm contexts."""
        L = N.lib()
        B = self.n0 - 1
        tag = C.c_longlong()
        bound, LBo, dfo = C.c_double(), C.c_double(), C.c_double()
        imp, st = C.c_int(), C.c_int()
        frac = np.zeros(B)
        xb = np.zeros(B, np.int32)
        stats = np.zeros(NSTATS, np.int32)
        if not L.bqb_ctx_pop(self.h, C.byref(tag), C.byref(bound), _d(frac), _i(xb), C.byref(LBo), C.byref(imp),
                             C.byref(dfo), C.byref(st), _i(stats)):
            return None
        nt, npn, nh = C.c_int(), C.c_int(), C.c_int()
        pt, pp, ph = N.iptr(), N.iptr(), N.iptr()
        pg = N.dptr()
        L.bqb_ctx_last_cuts(self.h, C.byref(nt), C.byref(pt), C.byref(npn), C.byref(pp), C.byref(nh), C.byref(ph),
                            C.byref(pg))
        m = nt.value + npn.value + nh.value
        cuts = dict(tri=np.ctypeslib.as_array(pt, (nt.value, 4)).copy() if nt.value else np.zeros((0, 4), np.int32),
                    pent=np.ctypeslib.as_array(pp, (npn.value, 8)).copy() if npn.value else np.zeros((0, 8), np.int32),
                    hept=np.ctypeslib.as_array(ph, (nh.value, 8)).copy() if nh.value else np.zeros((0, 8), np.int32),
                    gamma=np.ctypeslib.as_array(pg, (m,)).copy() if m else np.zeros(0))
        return dict(tag=tag.value, bound=bound.value, fracsol=frac, xbest=xb, LB=LBo.value, improved=imp.value,
                    diff=dfo.value, status=st.value, stats=stats, cuts=cuts)


if __name__ == "__main__":
    import sys
    if len(sys.argv) > 1 and sys.argv[1] == "copy-probe":
        nb = int(sys.argv[2]) if len(sys.argv) > 2 else 256 << 20
        ms = copy_probe(nb, 10)
        print(f"copy of {nb} bytes: {ms:.4f} ms, {2 * nb / ms / 1e6:.1f} GB/s")
