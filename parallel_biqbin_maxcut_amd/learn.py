"""The learned warm start and cut ranking: data from the solver's own runs,
the 3 x 256 MLP trained on the GPU by the library's kernels, and its use
in the node bound.

Data (solver-generated): every graph's root node is bounded with the
context recording the features of (a quarter of) the cuts each separation
appends, at the moment they are appended, and with trajectory capture on;
a recorded cut is labelled with its multiplier in the node's final cut set
(0 if it was purged before the end, src/cutting_planes.c:151, 336, 529).
Targets: log1p(gamma / 0.05) and [gamma > 1e-5].

The features of a cut (computed on the device, bqb_learn.h; cut_features
below restates them for the parity test):
   0 B_c(X)            1 B_c(X) - 1        2-4 one-hot cut size 3, 5, 7
   5 (type - 1) / k    6 mean |X_ab|       7 min s_ab X_ab     8 max s_ab X_ab
   9 mean s_ab L_ab / lscale               10 mean |L_ab| / lscale
  11 n / 128          12 round / 20       13 cuts / (10 n)    14 relative gap
  15 1                (16-31 zero)
with s_ab = -e_a e_b the cut's sign pattern (src/cutting_planes.c:20-39,
239-244, 431-438), lscale the mean |L_ij| over the node's off-diagonal, B_c
the cut operator (src/operators.c:51-132)."""
import ctypes as C
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import _native as N
from . import api, instances as I

NPARAM = 144144
NF = 32
GAMMA_SCALE = 0.05


def declare(L):
    vp = C.c_void_p
    fp = C.POINTER(C.c_float)
    L.bqb_mlp_create.argtypes = [fp, C.c_uint64]
    L.bqb_mlp_create.restype = vp
    L.bqb_mlp_destroy.argtypes = [vp]
    L.bqb_mlp_get_params.argtypes = [vp, fp]
    L.bqb_mlp_set_params.argtypes = [vp, fp]
    L.bqb_mlp_forward.argtypes = [vp, C.c_int, fp, fp]
    L.bqb_mlp_gradient.argtypes = [vp, C.c_int, fp, fp, fp, C.POINTER(C.c_double)]
    L.bqb_mlp_train.argtypes = [vp, C.c_int, fp, fp, C.c_int, C.c_int, C.c_float, C.c_uint64, N.dptr]
    L.bqb_cut_features.argtypes = [C.c_int, N.dptr, N.dptr, C.c_int, N.iptr, C.c_int, N.iptr, C.c_int, N.iptr,
                                   C.c_float, C.c_float, fp]
    L.bqb_ctx_set_model.argtypes = [vp, vp, C.c_float]
    L.bqb_ctx_learn_record.argtypes = [vp, C.c_int]
    L.bqb_ctx_learn_records.argtypes = [vp, N.iptr, fp, N.iptr]


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class MLP:
    """The cut model on the device (bqb_mlp_*)."""

    def __init__(self, params=None, seed=1):
        L = N.lib()
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float32)
        self.h = L.bqb_mlp_create(_f(p) if p is not None else None, seed)
        if not self.h:
            raise RuntimeError("bqb_mlp_create: " + L.bqb_last_error().decode())

    def close(self):
        if self.h:
            N.lib().bqb_mlp_destroy(self.h)
            self.h = None

    def params(self):
        out = np.zeros(NPARAM, dtype=np.float32)
        N.check(N.lib().bqb_mlp_get_params(self.h, _f(out)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        N.check(N.lib().bqb_mlp_set_params(self.h, _f(p)))

    def forward(self, feat):
        feat = np.ascontiguousarray(feat, dtype=np.float32)
        out = np.zeros((feat.shape[0], 2), dtype=np.float32)
        N.check(N.lib().bqb_mlp_forward(self.h, feat.shape[0], _f(feat), _f(out)))
        return out

    def gradient(self, feat, target):
        feat = np.ascontiguousarray(feat, dtype=np.float32)
        target = np.ascontiguousarray(target, dtype=np.float32)
        g = np.zeros(NPARAM, dtype=np.float32)
        loss = C.c_double()
        N.check(N.lib().bqb_mlp_gradient(self.h, feat.shape[0], _f(feat), _f(target), _f(g), C.byref(loss)))
        return g, loss.value

    def train(self, feat, target, epochs=4, batch=1024, lr=1e-3, seed=7):
        feat = np.ascontiguousarray(feat, dtype=np.float32)
        target = np.ascontiguousarray(target, dtype=np.float32)
        losses = np.zeros(epochs)
        N.check(N.lib().bqb_mlp_train(self.h, feat.shape[0], _f(feat), _f(target), epochs, batch, lr, seed,
                                      losses.ctypes.data_as(N.dptr)))
        return losses


def pad_features(f16):
    """[M][16] recorded features -> [M][32] model input."""
    out = np.zeros((f16.shape[0], NF), dtype=np.float32)
    out[:, :16] = f16
    return out


# ------------------------------------------------------------ the feature spec
def _sign_pattern(k, t):
    if k == 3:
        e = [1, 1, 1]
        if t == 2:
            e[2] = -1
        elif t == 3:
            e[1] = -1
        elif t == 4:
            e[0] = -1
        return e
    return [-1 if a < t - 1 else 1 for a in range(k)]


def cut_features(X, L, cuts, round_, gap):
    """Features of cuts [(k, type, vertices)] of one node (numpy, FP64)."""
    n = X.shape[0]
    off = ~np.eye(n, dtype=bool)
    lscale = np.abs(L[off]).mean() + 1e-12
    out = np.zeros((len(cuts), NF))
    for c, (k, t, p) in enumerate(cuts):
        e = _sign_pattern(k, t)
        s, x, l = [], [], []
        for a in range(k):
            for b in range(a + 1, k):
                s.append(-e[a] * e[b])
                x.append(X[p[a], p[b]])
                l.append(L[p[a], p[b]])
        s, x, l = np.array(s, float), np.array(x), np.array(l)
        coef = {3: 1.0, 5: 0.5, 7: 1.0 / 3.0}[k]
        bx = coef * np.sum(s * x)
        out[c, :16] = [bx, bx - 1.0, k == 3, k == 5, k == 7, (t - 1) / k, np.mean(np.abs(x)), np.min(s * x),
                       np.max(s * x), np.mean(s * l) / lscale, np.mean(np.abs(l)) / lscale, n / 128.0,
                       round_ / 20.0, len(cuts) / (10.0 * n), gap, 1.0]
    return out


def device_features(X, L, cuts, round_, gap):
    """The same features computed by the library's device code."""
    n = X.shape[0]
    tri = [(p[0], p[1], p[2], t) for k, t, p in cuts if k == 3]
    pent = [[t] + list(p[:5]) for k, t, p in cuts if k == 5]
    hept = [[t] + list(p[:7]) for k, t, p in cuts if k == 7]
    order = [c for c in cuts if c[0] == 3] + [c for c in cuts if c[0] == 5] + [c for c in cuts if c[0] == 7]
    ti = np.array(tri or [[0] * 4], dtype=np.int32)
    pi = np.array(pent or [[0] * 6], dtype=np.int32)
    hi = np.array(hept or [[0] * 8], dtype=np.int32)
    Xf = np.asfortranarray(X, dtype=np.float64)
    Lf = np.asfortranarray(L, dtype=np.float64)
    out = np.zeros((len(cuts), NF), dtype=np.float32)
    N.check(N.lib().bqb_cut_features(n, Xf.ctypes.data_as(N.dptr), Lf.ctypes.data_as(N.dptr), len(tri),
                                     ti.ctypes.data_as(N.iptr), len(pent), pi.ctypes.data_as(N.iptr), len(hept),
                                     hi.ctypes.data_as(N.iptr), round_, gap, _f(out)))
    return out, order


# ------------------------------------------------------------ training data
def _cut_keys_final(tr):
    """{(k, type, vertices...): gamma} of a trajectory's final cut set."""
    d = {}
    g = tr["gamma"]
    for c, (i, j, k, t) in enumerate(tr["tri"]):
        d[(3, int(t), int(i), int(j), int(k))] = g[c]
    o = len(tr["tri"])
    for c, row in enumerate(tr["pent"]):
        d[(5, int(row[0])) + tuple(int(v) for v in row[1:6])] = g[o + c]
    o += len(tr["pent"])
    for c, row in enumerate(tr["hept"]):
        d[(7, int(row[0])) + tuple(int(v) for v in row[1:8])] = g[o + c]
    return d


def _one_graph(args):
    adj, seed, params, cap = args
    SP = api.maxcut_L(adj)
    ctx = api.BoundContext(SP, params=params, max_nodes=1)
    try:
        L = N.lib()
        ctx.capture(True)
        N.check(L.bqb_ctx_learn_record(ctx.h, cap))
        z = np.zeros((1, SP.shape[0] - 1), np.int32)
        ctx.bound(z, z, LB=0.0, is_root=1, seeds=np.uint64(seed))
        cnt = C.c_int()
        feat = np.zeros((cap, 16), dtype=np.float32)
        key = np.zeros((cap, 12), dtype=np.int32)
        N.check(L.bqb_ctx_learn_records(ctx.h, C.byref(cnt), _f(feat), key.ctypes.data_as(N.iptr)))
        trs = ctx.trajectories()
    finally:
        ctx.close()
    k = cnt.value
    final = _cut_keys_final(trs[0]) if trs else {}
    gam = np.zeros(k)
    for r in range(k):
        kk = key[r]
        size = int(kk[0])
        kt = (size, int(kk[1])) + tuple(int(v) for v in kk[2:2 + size])
        gam[r] = final.get(kt, 0.0)
    return feat[:k], gam


def generate(graphs, params, seed=1, workers=8, cap=60000, progress=None):
    """Solver-generated training pairs from the root nodes of `graphs`
    (adjacency matrices): features [M][32] and targets [M][2].  progress:
    optional callable(done, total) after every block of graphs."""
    jobs = [(adj, seed + i, params, cap) for i, adj in enumerate(graphs)]
    res = []
    with ThreadPoolExecutor(max_workers=workers) as ex:
        for b in range(0, len(jobs), 100):
            res += list(ex.map(_one_graph, jobs[b:b + 100]))
            if progress:
                progress(len(res), len(jobs))
    feat = np.concatenate([r[0] for r in res]) if res else np.zeros((0, 16), np.float32)
    # multipliers are >= 0; the bundle's projection can leave -1e-17
    gam = np.maximum(np.concatenate([r[1] for r in res]), 0.0) if res else np.zeros(0)
    target = np.stack([np.log1p(gam / GAMMA_SCALE), (gam > 1e-5).astype(float)], axis=1).astype(np.float32)
    return pad_features(feat), target


def training_graphs(count, seed):
    """G(80, p ~ U[0.1, 0.9]) with weights uniform in {-1, 0, +1}."""
    rng = np.random.default_rng(seed)
    return [I.gnp_adj(80, float(rng.uniform(0.1, 0.9)), seed * 100003 + i, weights="zpm1") for i in range(count)]


# ------------------------------------------------------------ gain
def _bound_root(args):
    adj, params, seed, model, drop = args
    SP = api.maxcut_L(adj)
    ctx = api.BoundContext(SP, params=params, max_nodes=1)
    try:
        if model is not None:
            N.check(N.lib().bqb_ctx_set_model(ctx.h, model.h, drop))
        ctx.capture(True)
        z = np.zeros((1, SP.shape[0] - 1), np.int32)
        r = ctx.bound(z, z, LB=0.0, is_root=1, seeds=np.uint64(seed))
        tr = ctx.trajectories()
    finally:
        ctx.close()
    rb = tr[0]["round_bound"] if tr else np.zeros(0)
    return float(r["bound"][0]), float(r["LB"][0]), int(r["stats"][0][4]), int(r["stats"][0][3]), \
        int(r["stats"][0][5]), int(r["status"][0]), rb


def evaluate_gain(model, graphs, params, drop=0.0, seed=11, workers=8):
    """Root bounds of `graphs` with the plain cold-start bundle and with the
    model's warm start (and ranking when drop > 0), same parameters and
    seeds.  Returns per-graph arrays and a summary: the bound gap closed
    (bound - best cut found by either run), IPM iterations and calls."""
    jobs_p = [(adj, params, seed + i, None, 0.0) for i, adj in enumerate(graphs)]
    jobs_l = [(adj, params, seed + i, model, drop) for i, adj in enumerate(graphs)]
    with ThreadPoolExecutor(max_workers=workers) as ex:
        plain = list(ex.map(_bound_root, jobs_p))
        learned = list(ex.map(_bound_root, jobs_l))
    bp = np.array([r[0] for r in plain])
    bl = np.array([r[0] for r in learned])
    lb = np.maximum([r[1] for r in plain], [r[1] for r in learned])
    gp, gl = bp - lb, bl - lb
    ip = np.array([r[2] for r in plain])
    il = np.array([r[2] for r in learned])
    summary = {
        "graphs": len(graphs),
        "status_ok": bool(all(r[5] == 0 for r in plain + learned)),
        "mean_gap_plain": float(gp.mean()), "mean_gap_learned": float(gl.mean()),
        "gap_closed_rel": float((gp.mean() - gl.mean()) / max(gp.mean(), 1e-12)),
        "graphs_better": int(np.sum(bl < bp - 1e-9)), "graphs_worse": int(np.sum(bl > bp + 1e-9)),
        "ipm_iters_plain": int(ip.sum()), "ipm_iters_learned": int(il.sum()),
        "ipm_iters_ratio": float(il.sum() / max(ip.sum(), 1)),
        "max_cuts_plain": float(np.mean([r[4] for r in plain])), "max_cuts_learned": float(np.mean([r[4] for r in learned])),
    }
    # mean gap after each outer round (graphs that ran that many rounds in both)
    R = max(len(r[6]) for r in plain + learned)
    per_round = []
    for k in range(R):
        idx = [i for i in range(len(graphs)) if len(plain[i][6]) > k and len(learned[i][6]) > k]
        if not idx:
            break
        per_round.append([round(float(np.mean([plain[i][6][k] - lb[i] for i in idx])), 4),
                          round(float(np.mean([learned[i][6][k] - lb[i] for i in idx])), 4)])
    summary["gap_by_round_plain_learned"] = per_round
    return dict(bound_plain=bp, bound_learned=bl, lb=lb, iters_plain=ip, iters_learned=il), summary


def eval_graphs(count=100, seed=4242):
    """Held-out G(120, 0.5) with weights in {-1, +1}."""
    return [I.gnp_adj(120, 0.5, seed + i, weights="pm1") for i in range(count)]
