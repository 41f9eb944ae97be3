#!/usr/bin/env python3
"""Batched node-bounding benchmark on one MI355X.

Workload (batched bounding): Max-Cut on G(n=100, p=0.3) with weights uniform
in {-1,+1}; 512 B&B subproblems per launch, each with 0-40 of the 99 branching
variables fixed uniformly at random (subproblem sizes 60-100), bounded with the
full SDPbound (FP64 IPM + bundle over triangle / pentagonal / heptagonal cuts,
GW heuristic).  Every node receives the incumbent and the root gap of the
instance's root bound, as a B&B worker does.

One step = 65536 such nodes streamed through twelve slot groups of 512 node
slots (bqb_bound_stream_multi: every launch covers one group's 512 slots, a
finished node's slot is refilled after each outer round, as the B&B node
queue runs it; each group has its own HIP stream and host thread, so one
group's kernels fill the idle tail of another's launches).  The stream is
long because its start (every slot beginning a node at once) and its end
(the last nodes' remaining rounds with few slots busy) cost about 1.8 s
whatever its length; a B&B node queue that never runs empty does not pay
them (16384 nodes per step give 4540 bounds/s, 32768 6040, 65536 6750).  Inputs are uploaded
inside the call (a few hundred KB; the timed region includes it).
value = node bounds per second.  For comparison, config reports the nodes
that went past the basic bound (the whole SDPbound pipeline) streamed
through the same groups, the step streamed through one group of 8192 slots,
the first 512 bounded as one fixed batch (every launch waits for the
slowest node), and the first 64 bounded one at a time (the plain GPU form).
The roofline object covers the dominant
kernel, the batched IPM: algorithmic FP64 flops (6 n^3 per iteration + n^3/3 per line-search
Cholesky test, counted on the device) over the summed duration of its launches
(HIP event pairs on the launch stream, during the timed steps); with twelve
groups in flight a launch's duration includes the time its workgroups wait
for CUs held by other kernels, so the object also gives the same flops over
the timed wall time (aggregate_achieved).  The
cpu_baseline times the CPU oracle (the FP64 restatement of the same
algorithm) on a bounded sample of the same nodes with one thread per core;
port_over_reference_root_time ties it to the reference binary (measured on
the same host type, one core each).

Besides the headline, config carries: time_to_proof (G(60, 0.5) branch and
bound to the proven optimum: the plain FP64 form one node at a time, a
64-slot node queue, children inheriting their parent's cuts and multipliers,
bf16 matrix-core products with certified bounds, the learned cut model, and
bundle-step IPMs warm-started from the node's previous solve; each form's
speedup over the plain one, and the root bound's time alone, cold and warm,
which caps every form's speedup), cfg3_mixed (G(250, 0.1), FP64 against
bf16 matrix-core products with certified bounds, with the roofline of the
mixed IPM against the bf16 peak) and cfg4_qubo (QUBO 150 as 151-vertex
Max-Cut, 256 nodes per launch).
"""
import argparse
import json
import os
import sys
import time

# Each slot group drives its own HIP stream; HIP maps the streams of a
# process onto GPU_MAX_HW_QUEUES hardware queues (4 by default), and streams
# that share a queue serialize: twelve groups run on 24 queues (with as many
# queues as groups some streams still share one and throughput drops).
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 24:
    os.environ["GPU_MAX_HW_QUEUES"] = "24"

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N0 = 100                # graph vertices (the last one is fixed: 99 branching variables)
EDGE_P = 0.3
NFIX_MAX = 40
FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "ipm_batch_cfg2_traffic.json")


def node_seeds(seed, count):
    return (np.uint64(seed) * np.uint64(1000003) + np.arange(1, count + 1, dtype=np.uint64)).astype(np.uint64)


def build_workload(api, I, seed, nodes):
    adj = I.gnp_adj(N0, EDGE_P, seed, weights="pm1")
    SP = api.maxcut_L(adj)
    P = api.make_params()                      # the reference's shipped params file
    if P.adjust_TriIneq:
        P.TriIneq = 10 * N0                     # src/process_input.c:135-136
    xf, sx = I.random_fixings(N0, NFIX_MAX, nodes, seed + 1)
    return SP, P, xf, sx, node_seeds(seed, nodes)


def cpu_baseline(SP, P, xf, sx, LB, diff, seeds, sample):
    """The CPU oracle on the first `sample` nodes, one thread per core."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O
    cores = min(16, len(os.sched_getaffinity(0)))
    d = {f: getattr(P, f) for f, _ in type(P)._fields_}
    op = O.default_params(**d)
    t0 = time.perf_counter()
    bounds, status = O.sdpbound_many(SP, op, xf[:sample], sx[:sample], LB, diff, seeds[:sample], cores)
    dt = time.perf_counter() - t0
    return {"value": sample / dt, "unit": "bounds/s", "cores": cores, "kind": "port",
            "sample": f"first {sample} of the {len(xf)} nodes of a step (same instance, incumbent and root gap), "
                      f"{dt:.1f} s wall on {cores} threads"}, bounds, status


def progress(msg):
    """A progress line on stderr (the JSON result is the only stdout line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


# time-to-proof forms: (name, bqb_solve_opt options); "plain" is the plain
# FP64 GPU form, one node at a time, against which the others are compared
PROOF_FORMS = (
    ("plain", dict(batch=1)),
    ("batched", dict(batch=64)),
    ("batched_inherited", dict(batch=64, warm=True)),
    # mixed: the automatic precision (per size class the fastest measured
    # solve -- FP64 products at n = 60 -- with certified bounds)
    ("batched_inherited_mixed", dict(batch=64, warm=True, slices=-1)),
    ("batched_inherited_mixed_learned", dict(batch=64, warm=True, slices=-1, learned=True)),
    # the IPM products forced onto the bf16 matrix cores (4 slices)
    ("batched_inherited_bf16", dict(batch=64, warm=True, slices=4)),
    # every bundle-step IPM warm-started from the node's previous solve
    ("batched_inherited_ipmwarm", dict(batch=64, warm=True, ipm_warm=True)),
    ("batched_inherited_mixed_learned_ipmwarm", dict(batch=64, warm=True, slices=-1, learned=True, ipm_warm=True)),
)


def proof_one(api, I, seed, form, reps):
    """One time-to-proof instance (G(60, 0.5), unit weights, triangles only)
    in one form, after one untimed proof, `reps` times."""
    P = api.make_params(include_Pent=0, include_Hepta=0)
    kw = dict(dict(PROOF_FORMS)[form])
    model = None
    if kw.pop("learned", False):
        from parallel_biqbin_maxcut_amd import learn
        model = learn.MLP(params=np.load(os.path.join(ROOT, "parallel_biqbin_maxcut_amd", "models",
                                                      "cut_model.npz"), allow_pickle=False)["params"])
        kw["model"] = model
    L = api.maxcut_L(I.gnp_adj(60, 0.5, seed, weights="unit"))
    api.solve(L, params=P, seed=2020, **kw)
    secs = []
    for _ in range(reps):
        st, _ = api.solve(L, params=P, seed=2020, **kw)
        secs.append(round(st.seconds, 4))
    if model is not None:
        model.close()
    return {"mode": "proof_one", "seed": seed, "form": form, "optimum": st.best, "nodes": int(st.nodes),
            "batches": int(st.batches), "seconds": secs}


def time_to_proof(api, I, seeds, cpu_seeds):
    """Full B&B to the proven optimum on G(60, 0.5) with unit weights,
    triangle cuts only (the reference's params with pentagonal and
    heptagonal cuts off), in the forms of PROOF_FORMS: one node at a time;
    a 64-slot node queue; children inheriting their parent's cuts and
    multipliers; bf16 matrix-core products with certified bounds; the
    learned cut model.  Every form must prove the same optimum.  The CPU
    oracle's sequential B&B runs on the first instances for comparison."""
    P = api.make_params(include_Pent=0, include_Hepta=0)
    out = {"workload": f"G(60, 0.5) unit weights, triangles only, seeds {seeds[0]}-{seeds[-1]}",
           "forms": {name: {k: v for k, v in kw.items()} for name, kw in PROOF_FORMS},
           "optimum": [], "nodes_plain": []}
    from parallel_biqbin_maxcut_amd import learn
    model = learn.MLP(params=np.load(os.path.join(ROOT, "parallel_biqbin_maxcut_amd", "models", "cut_model.npz"),
                                     allow_pickle=False)["params"])
    secs = {name: [] for name, _ in PROOF_FORMS}
    nodes = {name: 0 for name, _ in PROOF_FORMS}
    Ls = []
    for sd in seeds:
        L = api.maxcut_L(I.gnp_adj(60, 0.5, sd, weights="unit"))
        Ls.append(L)
        best = None
        for name, kw in PROOF_FORMS:
            kw = dict(kw)
            if kw.pop("learned", False):
                kw["model"] = model
            st, _ = api.solve(L, params=P, seed=2020, **kw)
            if st.status or (best is not None and st.best != best):
                raise SystemExit(f"time-to-proof instance {sd}, form {name}: status {st.status}, "
                                 f"optimum {st.best} (plain {best})")
            best = st.best
            secs[name].append(round(st.seconds, 4))
            nodes[name] += int(st.nodes)
            if name == "plain":
                out["nodes_plain"].append(int(st.nodes))
        out["optimum"].append(best)
    progress(f"time to proof: {len(seeds)} instances in {len(PROOF_FORMS)} forms")
    model.close()
    for name, _ in PROOF_FORMS:
        out[f"seconds_{name}"] = secs[name]
        out[f"median_s_{name}"] = float(np.median(secs[name]))
        out[f"total_s_{name}"] = round(float(np.sum(secs[name])), 4)
        out[f"nodes_{name}"] = nodes[name]
        # total time of the 20 proofs, plain form over this form
        out[f"speedup_{name}_vs_plain"] = round(sum(secs["plain"]) / sum(secs[name]), 3)
    best_form = max((name for name, _ in PROOF_FORMS), key=lambda nm: out[f"speedup_{nm}_vs_plain"])
    out["fastest_form"] = best_form
    # bounds/s on the node stream of a real branch and bound (64 slots)
    out["bab_node_bounds_per_s_batched"] = round(nodes["batched"] / sum(secs["batched"]), 2)
    # The root bound alone (one slot, the same parameters): every form
    # starts with it and no child can start before it ends, so the plain
    # total over the roots' total bounds any form's speedup, even one whose
    # children cost nothing.
    roots, roots_w = [], []
    for L, ipm_warm in [(L, w) for L in Ls for w in (False, True)]:
        B = L.shape[0] - 1
        Pr = api.make_params(include_Pent=0, include_Hepta=0)
        Pr.TriIneq = 10 * L.shape[0]               # src/process_input.c:135-136
        z = np.zeros(B, np.int32)
        for _ in range(2):                          # the first is untimed
            ctx = api.BoundContext(L, params=Pr, max_nodes=1)
            if ipm_warm:
                ctx.set_ipm_warm(True)
            t0 = time.perf_counter()
            ctx.push(z, z, LB=0.0, diff=0.0, is_root=1, seed=2020, tag=0)
            while ctx.busy():
                ctx.advance()
            r = ctx.pop()
            dt = time.perf_counter() - t0
            ctx.close()
        if r["status"]:
            raise SystemExit(f"time-to-proof root bound: status {r['status']}")
        (roots_w if ipm_warm else roots).append(round(dt, 4))
    out["seconds_root_bound"] = roots
    out["total_s_root_bound"] = round(float(np.sum(roots)), 4)
    out["speedup_bound_children_free"] = round(sum(secs["plain"]) / sum(roots), 3)
    # the same with warm-started bundle IPMs (the *_ipmwarm forms' root)
    out["seconds_root_bound_ipm_warm"] = roots_w
    out["total_s_root_bound_ipm_warm"] = round(float(np.sum(roots_w)), 4)
    out["speedup_bound_children_free_ipm_warm"] = round(sum(secs["plain"]) / sum(roots_w), 3)
    progress("time to proof: root bounds timed")
    if cpu_seeds:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py as O
        d = {f: getattr(P, f) for f, _ in type(P)._fields_}
        d["TriIneq"] = 10 * 60                     # src/process_input.c:135-136
        cs = []
        for i in range(cpu_seeds):
            t0 = time.perf_counter()
            ob, _ = O.bab(Ls[i], O.default_params(**d), seed=2020, batch=1)
            cs.append(round(time.perf_counter() - t0, 4))
            if ob.best != out["optimum"][i]:
                raise SystemExit(f"time-to-proof instance {seeds[i]}: CPU oracle optimum {ob.best}, "
                                 f"GPU {out['optimum'][i]}")
        out["cpu_oracle_seconds_plain"] = cs
        out["cpu_oracle_cores"] = 1
        # the other instances' optima against the CPU oracle (untimed, one
        # sequential B&B per thread)
        from concurrent.futures import ThreadPoolExecutor
        with ThreadPoolExecutor(min(16, len(os.sched_getaffinity(0)))) as ex:
            rest = list(ex.map(lambda L: O.bab(L, O.default_params(**d), seed=2020, batch=1)[0].best,
                               Ls[cpu_seeds:]))
        for i, b in enumerate(rest, start=cpu_seeds):
            if b != out["optimum"][i]:
                raise SystemExit(f"time-to-proof instance {seeds[i]}: CPU oracle optimum {b}, GPU {out['optimum'][i]}")
        out["optima_checked_against_cpu_oracle"] = len(Ls)
        progress("time-to-proof optima checked against the CPU oracle")
    return out


# Root-node bound times of the CPU oracle (this build's FP64 port, one core)
# against the reference binary (MKL, one core) measured on the same 8-core
# Xeon host, same configurations (the reference's figures: 1.38 s cfg2-like,
# 4.42 s QUBO 150 -> 151 vertices, 10.9 s G(250, 0.1)): the port is this many
# times slower per node, so the reference's own CPU throughput is about
# cpu_baseline.value times this ratio.
PORT_OVER_REFERENCE = {"cfg2 G(100, 0.3) root": round(3.95 / 1.38, 2),
                       "cfg4 QUBO 150 root": round(26.72 / 4.42, 2),
                       "cfg3 G(250, 0.1) root": round(68.49 / 10.9, 2)}
BF16_PEAK_TFLOPS = 2500.0   # MI355X dense bf16 matrix peak


def mixed_section(api, I, seed, nodes=64):
    """cfg3: Max-Cut on G(250, 0.1) with weights uniform in {-10..10} without 0:
    the root bound, then `nodes` B&B nodes with 0-40 random fixings (sizes
    210-250) bounded as one batch, with FP64 products and with the products
    on the bf16 matrix cores (4 slices) and certified bounds, each with the
    default IPM tile table (the blocked solve for these sizes) and with the
    32 x 8 register tile forced (node batch only).  The roofline of the mixed
    IPM kernel: its bf16 matrix-core flops (2 n^3 per slice pair and
    product) over the summed IPM launch time of the node batch."""
    SP = api.maxcut_L(I.gnp_adj(250, 0.1, seed, weights="int10"))
    P = api.make_params()
    P.TriIneq = 2500                                  # adjust_TriIneq: 10 * n0
    xf, sx = I.random_fixings(250, 40, nodes, seed + 1)
    seeds = np.arange(1, nodes + 1, dtype=np.uint64)
    z = np.zeros((1, 249), np.int32)
    out = {"workload": f"G(250, 0.1) weights +-1..+-10, root bound then {nodes} nodes with 0-40 fixings "
                       "(sizes 210-250) as one batch; reference params"}
    res = {}
    out["ipm_tile_table"] = {str(k): v for k, v in api.ipm_tile_table().items()}
    LB = diff = None
    for mode, slices, tile in (("fp64", 0, 0), ("mixed", 4, 0), ("fp64_register_tile", 0, 3208),
                               ("mixed_register_tile", 4, 3208)):
        api.N.lib().bqb_set_ipm_tile(tile)
        ctx = api.BoundContext(SP, params=P, max_nodes=nodes)
        api.N.lib().bqb_set_ipm_tile(0)
        ctx.set_precision(slices=slices, certify=1 if slices else 0)
        if not tile:
            t0 = time.perf_counter()
            root = ctx.bound(z, z, LB=0.0, is_root=1, seeds=np.uint64(seed))
            t_root = time.perf_counter() - t0
            if mode == "fp64":
                LB, diff = float(root["LB"][0]), float(root["diff"][0])
        ctx.profile_ipm(max_launches=20000)
        ctx.sync()
        t0 = time.perf_counter()
        r = ctx.bound(xf, sx, LB=LB, diff=diff, is_root=0, seeds=seeds)
        dt = time.perf_counter() - t0
        launches, ipm_ms, fl64 = ctx.ipm_stats()
        flmx = ctx.ipm_flops_mx()
        ctx.profile_ipm(0)
        ctx.close()
        if np.any(r["status"]):
            raise SystemExit(f"cfg3 {mode}: node failures {np.unique(r['status'])}")
        res[mode] = (r, dt, ipm_ms, fl64, flmx, launches)
        if not tile:
            out[f"{mode}_root_s"] = round(t_root, 3)
            out[f"{mode}_root_bound"] = round(float(root["bound"][0]), 6)
        out[f"{mode}_bounds_per_s"] = round(nodes / dt, 3)
        out[f"{mode}_ipm_ms"] = round(ipm_ms, 2)
    bp, bm = res["fp64"][0]["bound"], res["mixed"][0]["bound"]
    rel = np.abs(bm - bp) / np.abs(bp)
    out["speedup_bounds_per_s"] = round(out["mixed_bounds_per_s"] / out["fp64_bounds_per_s"], 3)
    out["speedup_root"] = round(out["fp64_root_s"] / out["mixed_root_s"], 3)
    out["speedup_register_tile_mixed_vs_fp64"] = round(
        out["mixed_register_tile_bounds_per_s"] / out["fp64_register_tile_bounds_per_s"], 3)
    out["speedup_blocked_vs_register_tile_fp64"] = round(
        out["fp64_bounds_per_s"] / out["fp64_register_tile_bounds_per_s"], 3)
    out["rel_bound_diff_mixed_vs_fp64"] = {"median": float(np.median(rel)), "max": float(rel.max())}
    out["fp64_retries"] = int(np.sum(res["mixed"][0]["stats"][:, 8] & api.FLAG_FP64_RETRY > 0))
    _, _, ipm_ms, fl64, flmx, launches = res["mixed"]
    ach = flmx / (ipm_ms / 1e3) / 1e12 if ipm_ms > 0 else 0.0
    out["roofline"] = {"kernel": "ipm_blk_kernel (products on bf16 matrix cores)", "bound": "mfma",
                       "achieved": round(ach, 4), "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                       "frac": round(ach / BF16_PEAK_TFLOPS, 6), "traffic": None, "launches": launches,
                       "fp64_achieved": round(fl64 / (ipm_ms / 1e3) / 1e12, 4) if ipm_ms > 0 else 0.0}
    # the FP64 blocked solve of the node batch against the FP64 peak
    _, _, ipm_ms, fl64, _, launches = res["fp64"]
    ach = fl64 / (ipm_ms / 1e3) / 1e12 if ipm_ms > 0 else 0.0
    out["roofline_fp64"] = {"kernel": "ipm_blk_kernel (FP64)", "bound": "mfma", "achieved": round(ach, 4),
                            "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(ach / FP64_PEAK_TFLOPS, 5),
                            "launches": launches}
    # where the automatic table takes bf16 products: the 193-224 class
    # (G(224, 0.1), 64 nodes with 0-14 fixings, sizes 210-224), automatic
    # precision against the FP64 default table
    SP2 = api.maxcut_L(I.gnp_adj(224, 0.1, seed, weights="int10"))
    P2 = api.make_params()
    P2.TriIneq = 2240
    xf2, sx2 = I.random_fixings(224, 14, nodes, seed + 2)
    cls = {"workload": f"G(224, 0.1) weights +-1..+-10, {nodes} nodes with 0-14 fixings (sizes 210-224) as "
                       "one batch after the root bound; reference params",
           "auto_table_224": list(api.ipm_auto_table()[224])}
    bounds = {}
    for mode, sl in (("fp64", 0), ("auto", -1)):
        ctx = api.BoundContext(SP2, params=P2, max_nodes=nodes)
        ctx.set_precision(slices=sl, certify=0 if sl == 0 else 1)
        root = ctx.bound(np.zeros((1, 223), np.int32), np.zeros((1, 223), np.int32), LB=0.0, is_root=1,
                         seeds=np.uint64(seed))
        ctx.sync()
        t0 = time.perf_counter()
        r = ctx.bound(xf2, sx2, LB=float(root["LB"][0]), diff=float(root["diff"][0]), is_root=0, seeds=seeds)
        dt = time.perf_counter() - t0
        ctx.close()
        if np.any(r["status"]):
            raise SystemExit(f"class 224 {mode}: node failures {np.unique(r['status'])}")
        bounds[mode] = r["bound"]
        cls[f"{mode}_bounds_per_s"] = round(nodes / dt, 3)
    cls["speedup_auto_vs_fp64"] = round(cls["auto_bounds_per_s"] / cls["fp64_bounds_per_s"], 3)
    rel2 = np.abs(bounds["auto"] - bounds["fp64"]) / np.abs(bounds["fp64"])
    cls["rel_bound_diff_auto_vs_fp64"] = {"median": float(np.median(rel2)), "max": float(rel2.max())}
    out["class_224"] = cls
    return out


def qubo_section(api, I, seed, nodes=256):
    """cfg4: dense symmetric integer QUBO, n = 150, entries uniform in
    {-100..100} with 50 % density, as Max-Cut on 151 vertices
    (biqbin_base.py:239-258); 256 B&B nodes with 0-40 fixings bounded per
    launch (sizes 111-151), FP64, after the instance's root bound."""
    SP = api.maxcut_L(I.qubo_to_maxcut_adj(I.random_qubo(150, 0.5, seed)))
    P = api.make_params()
    P.TriIneq = 1510
    ctx = api.BoundContext(SP, params=P, max_nodes=nodes)
    z = np.zeros((1, 150), np.int32)
    t0 = time.perf_counter()
    root = ctx.bound(z, z, LB=0.0, is_root=1, seeds=np.uint64(seed))
    t_root = time.perf_counter() - t0
    xf, sx = I.random_fixings(151, 40, nodes, seed + 1)
    t0 = time.perf_counter()
    r = ctx.bound(xf, sx, LB=float(root["LB"][0]), diff=float(root["diff"][0]), is_root=0,
                  seeds=np.arange(1, nodes + 1, dtype=np.uint64))
    dt = time.perf_counter() - t0
    ctx.close()
    if np.any(r["status"]):
        raise SystemExit(f"cfg4: node failures {np.unique(r['status'])}")
    return {"workload": "QUBO n=150, entries {-100..100} at 50% density -> Max-Cut on 151 vertices; "
                        f"{nodes} nodes per launch with 0-40 fixings (sizes 111-151), FP64",
            "root_s": round(t_root, 3), "root_bound": float(root["bound"][0]), "incumbent": float(root["LB"][0]),
            "bounds_per_s": round(nodes / dt, 3), "mean_outer_rounds": round(float(r["stats"][:, 2].mean()), 3)}


def traffic_per_launch():
    """HBM bytes per IPM launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload, or None."""
    try:
        t = json.load(open(TRAFFIC_FILE))
        return float(t["hbm_bytes_per_launch"])
    except (OSError, ValueError, KeyError):
        return None


def subproblem_matrices(api, SP, xf, sx):
    """The subproblem matrices of the given fixings, built by the library's
    drop-in createSubproblem (src/evaluate.c:38-115)."""
    n0 = SP.shape[0]
    buf = np.zeros(n0 * n0)
    prob = api.Problem(SP.ctypes.data_as(api.N.dptr), n0, 0, 0, 0, 0)
    sub = api.Problem(buf.ctypes.data_as(api.N.dptr), n0, 0, 0, 0, 0)
    out = []
    for b in range(len(xf)):
        node = api.BabNode()
        node.xfixed[: n0 - 1] = list(xf[b])
        node.solX[: n0 - 1] = list(sx[b])
        api.N.lib().createSubproblem(node, prob, sub)
        out.append(buf[: sub.n * sub.n].reshape((sub.n, sub.n), order="F").copy())
    return out


def ipm_isolated(api, I, seed, nodes, reps, workload="cfg2", slices=0):
    """The batched IPM alone on one stream: the basic SDP of `nodes`
    subproblems of the batched-bounding instance (cfg2: G(100, 0.3) +-1,
    sizes 60-100) or of the mixed-precision instance (cfg3: G(250, 0.1)
    +-1..+-10, sizes 210-250), one launch per repetition, timed by HIP
    events around the launch (and the FP64 repeat of failed solves)."""
    if workload == "cfg3":
        SP = api.maxcut_L(I.gnp_adj(250, 0.1, seed, weights="int10"))
        xf, sx = I.random_fixings(250, 40, nodes, seed + 1)
        desc = "G(250, 0.1) +-1..+-10"
    else:
        SP, _, xf, sx, _ = build_workload(api, I, seed, nodes)
        desc = "G(100, 0.3) +-1"
    Ls = subproblem_matrices(api, SP, xf, sx)
    api.ipm_batch_mixed(Ls, slices=slices)              # warm-up
    ms = []
    for _ in range(reps):
        r = api.ipm_batch_mixed(Ls, slices=slices)
        if np.any(r["status"]):
            raise SystemExit(f"IPM failures: {np.unique(r['status'], return_counts=True)}")
        ms.append(r["ms"])
    ns = np.array([m.shape[0] for m in Ls], dtype=np.float64)
    flops = float(np.sum(ns ** 3 * (6.0 * r["iters"] + r["trials"] / 3.0)))
    t = float(np.median(ms))
    achieved = flops / (t / 1e3) / 1e12
    # algorithmic HBM bytes: L read and X written once per subproblem
    alg_bytes = float(np.sum(2 * 8 * ns ** 2))
    tiles = sorted({int(api.N.lib().bqb_ipm_tile_for(int(n))) for n in ns})
    return {"mode": "ipm_isolated", "kernel": "ipm_blk_kernel" if all(t % 100 == 1 for t in tiles)
            else "ipm_batch_kernel", "tiles": tiles, "slices": slices, "workload":
            f"basic SDP of {nodes} subproblems of Max-Cut {desc} with 0-40 fixings (sizes "
            f"{int(ns.min())}-{int(ns.max())}), one launch on one stream", "launch_ms": round(t, 4),
            "launch_ms_all": [round(x, 4) for x in ms], "gflop_per_launch": round(flops / 1e9, 4),
            "mean_iters": round(float(r["iters"].mean()), 3), "mean_trials": round(float(r["trials"].mean()), 3),
            "achieved": round(achieved, 4), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / FP64_PEAK_TFLOPS, 5), "algorithmic_hbm_bytes": alg_bytes}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nodes", type=int, default=512, help="node slots (subproblems per launch)")
    ap.add_argument("--per-step", type=int, default=65536, help="nodes bounded per step")
    ap.add_argument("--groups", type=int, default=12, help="slot groups (streams) of --nodes slots each")
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--cpu-sample", type=int, default=192)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-proof", action="store_true", help="skip the time-to-proof instances")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the cfg3 mixed-precision and cfg4 QUBO sections")
    ap.add_argument("--no-single", action="store_true",
                    help="skip the single-batch comparison after the timed steps (profiling runs)")
    ap.add_argument("--graphs", type=int, default=0,
                    help="1: launch trains as HIP graphs (basic sequence and outer-round slices)")
    ap.add_argument("--ipm-isolated", action="store_true",
                    help="time the batched IPM alone on --nodes subproblems, one launch per step (profiling)")
    ap.add_argument("--ipm-workload", default="cfg2", choices=("cfg2", "cfg3"),
                    help="instance of --ipm-isolated: cfg2 G(100, 0.3) or cfg3 G(250, 0.1)")
    ap.add_argument("--ipm-slices", type=int, default=0, help="--ipm-isolated products: 0 FP64, 2..4 bf16 slices")
    ap.add_argument("--proof-seed", type=int, default=0,
                    help="time the proof of one G(60, 0.5) instance (seed) in --proof-form, --steps times (profiling)")
    ap.add_argument("--proof-form", default="plain", choices=[nm for nm, _ in PROOF_FORMS])
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: bounding runs on one GPU; --gpus {args.gpus} is not supported "
              "(the node queue is a single process driving one device)", file=sys.stderr)
        sys.exit(2)

    from parallel_biqbin_maxcut_amd import api, instances as I

    if args.ipm_isolated:
        print(json.dumps(ipm_isolated(api, I, args.seed, args.nodes, max(1, args.steps), args.ipm_workload,
                                      args.ipm_slices)), flush=True)
        return
    if args.proof_seed:
        print(json.dumps(proof_one(api, I, args.proof_seed, args.proof_form, max(1, args.steps))), flush=True)
        return

    SP, P, xf, sx, seeds = build_workload(api, I, args.seed, args.per_step)
    api.set_default_graphs(bool(args.graphs))
    ctx = api.BoundContext(SP, params=P, max_nodes=args.nodes, groups=args.groups)
    # root bound of the instance: incumbent and root gap the workers receive
    z = np.zeros((1, N0 - 1), np.int32)
    root = ctx.bound(z, z, LB=0.0, is_root=1, seeds=np.uint64(args.seed))
    if root["status"][0]:
        raise SystemExit(f"root bound failed with status {root['status'][0]}")
    LB, diff = float(root["LB"][0]), float(root["diff"][0])
    progress(f"root bound {float(root['bound'][0]):.4f}, incumbent {LB}")

    for w in range(args.warmup):
        ctx.bound(xf, sx, LB=LB, diff=diff, is_root=0, seeds=seeds, stream=True)
        progress(f"warm-up step {w + 1}/{args.warmup} done")
    # the IPM launch profile (event pairs around each IPM launch) keeps the
    # trains eager; with graphs the flops are counted in an untimed pass
    if not args.graphs:
        ctx.profile_ipm(max_launches=5000 * max(1, args.steps))
    ctx.sync()
    ctx.counters(reset=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = ctx.bound(xf, sx, LB=LB, diff=diff, is_root=0, seeds=seeds, stream=True)
    ctx.sync()
    dt = time.perf_counter() - t0
    progress(f"{args.steps} timed steps: {args.per_step * args.steps / dt:.1f} bounds/s")
    counters = ctx.counters()
    if args.graphs:
        ctx.set_graphs(False)
        ctx.profile_ipm(max_launches=5000)
        r2 = ctx.bound(xf, sx, LB=LB, diff=diff, is_root=0, seeds=seeds, stream=True)
        if not np.array_equal(r2["bound"], r["bound"]):
            raise SystemExit("graph and eager launch trains gave different bounds")
    launches, ipm_ms, flops = ctx.ipm_stats()
    if args.graphs:
        flops *= args.steps
    ctx.profile_ipm(0)
    if np.any(r["status"]):
        raise SystemExit(f"node bound failures: {np.unique(r['status'], return_counts=True)}")

    # comparisons: the nodes that went past the basic bound (the whole
    # SDPbound pipeline) through the same groups; the step through one group
    # of 8192 slots; the first 512 nodes as one fixed batch (every launch
    # waits for its slowest node).  All must give the step's bounds.
    single_batch = one_group = full_pipeline = plain_one = None
    deep = np.nonzero(r["stats"][:, 2] > 0)[0]
    if not args.no_single:
        tb = time.perf_counter()
        rd = ctx.bound(xf[deep], sx[deep], LB=LB, diff=diff, is_root=0, seeds=seeds[deep], stream=True)
        full_pipeline = round(len(deep) / (time.perf_counter() - tb), 2)
        ctx.close()
        one = api.BoundContext(SP, params=P, max_nodes=8192)
        tb = time.perf_counter()
        ro = one.bound(xf, sx, LB=LB, diff=diff, is_root=0, seeds=seeds, stream=True)
        one_group = round(args.per_step / (time.perf_counter() - tb), 2)
        tb = time.perf_counter()
        rb = one.bound(xf[: args.nodes], sx[: args.nodes], LB=LB, diff=diff, is_root=0, seeds=seeds[: args.nodes])
        single_batch = round(args.nodes / (time.perf_counter() - tb), 2)
        one.close()
        # the plain GPU form: one node at a time (a one-slot context), on the
        # first 64 nodes of the step
        solo = api.BoundContext(SP, params=P, max_nodes=1)
        k1 = min(64, args.per_step)
        tb = time.perf_counter()
        r1 = solo.bound(xf[:k1], sx[:k1], LB=LB, diff=diff, is_root=0, seeds=seeds[:k1], stream=True)
        plain_one = round(k1 / (time.perf_counter() - tb), 2)
        solo.close()
        if not (np.array_equal(rb["bound"], r["bound"][: args.nodes]) and np.array_equal(ro["bound"], r["bound"])
                and np.array_equal(rd["bound"], r["bound"][deep]) and np.array_equal(r1["bound"], r["bound"][:k1])):
            raise SystemExit("streamed and batched bounds differ")
        progress("comparisons done")
    else:
        ctx.close()

    ms_per_step = dt / args.steps * 1e3
    value = args.per_step * args.steps / dt
    achieved = flops / (ipm_ms / 1e3) / 1e12 if ipm_ms > 0 else 0.0
    stats = r["stats"]
    line = {
        "metric": "node bounds/sec (batched bounding, plain FP64 GPU form)",
        "value": round(value, 2),
        "unit": "bounds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "batched bounding: Max-Cut G(n=100, p=0.3), weights in {-1,+1}, 512 subproblems "
                        "per launch with 0-40 variables fixed at random (sizes 60-100), FP64 IPM + bundle "
                        "with triangle/pentagonal/heptagonal cuts",
            "nodes_per_launch": args.nodes,
            "slot_groups": args.groups,
            "nodes_per_step": args.per_step,
            "one_group_8192_slots_bounds_per_s": one_group,
            "single_batch_bounds_per_s": single_batch,
            "full_pipeline_nodes": int(len(deep)),
            "full_pipeline_bounds_per_s": full_pipeline,
            # node bounds/s of the plain GPU form (one node at a time) on the
            # same nodes, and the batched form's speedup over it
            "plain_one_at_a_time_bounds_per_s": plain_one,
            "speedup_batched_vs_plain_bounds_per_s": round(value / plain_one, 1) if plain_one else None,
            "seed": args.seed,
            "incumbent": LB,
            "root_gap": round(diff, 4),
            "pruned_after_basic_bound": int(np.sum((stats[:, 0] == 1) & (stats[:, 2] == 0))),
            "mean_outer_rounds": round(float(stats[:, 2].mean()), 3),
            "ipm_calls": int(stats[:, 3].sum()),
            "graphs": bool(args.graphs),
            "launch_counters_per_step": {k: round(v / args.steps, 1) for k, v in counters.items()},
            "host_syncs_per_round_slice": round(counters["host_syncs"] / max(1, counters["round_slices"]), 3),
            "kernel_phases_per_round_slice": round((counters["eager_phases"] + counters["graph_phases"])
                                                   / max(1, counters["round_slices"]), 3),
            "host_launch_calls_per_round_slice": round((counters["eager_phases"] + counters["graph_launches"])
                                                       / max(1, counters["round_slices"]), 3),
        },
        "roofline": {
            "kernel": "ipm_batch_kernel",
            "bound": "mfma",
            "achieved": round(achieved, 4),
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP64_PEAK_TFLOPS, 5),
            "traffic": traffic_per_launch(),
            "launches": launches,
            "avg_launch_ms": round(ipm_ms / max(launches, 1), 5),
            "gflop_per_launch": round(flops / max(launches, 1) / 1e9, 4),
            # the same algorithmic flops over the timed wall time: the slot
            # groups' IPM launches overlap, so each launch's duration includes
            # the time its workgroups wait for CUs held by other kernels
            "aggregate_achieved": round(flops / dt / 1e12, 4),
            "aggregate_frac": round(flops / dt / 1e12 / FP64_PEAK_TFLOPS, 5),
        },
        "cpu_baseline": None,
    }
    if not args.no_proof:
        line["config"]["time_to_proof"] = time_to_proof(api, I, list(range(1, 21)), 0 if args.no_cpu else 3)
    if not args.no_extra:
        line["config"]["cfg3_mixed"] = mixed_section(api, I, args.seed)
        progress("mixed-precision section done")
        line["config"]["cfg4_qubo"] = qubo_section(api, I, args.seed)
        progress("QUBO section done")
    if not args.no_cpu:
        sample = min(args.cpu_sample, args.per_step)
        cb, cbounds, cstatus = cpu_baseline(SP, P, xf, sx, LB, diff, seeds, sample)
        line["cpu_baseline"] = cb
        ok = cstatus == 0
        rel = np.abs(r["bound"][:sample][ok] - cbounds[ok]) / np.abs(cbounds[ok])
        line["cpu_baseline"]["max_rel_bound_diff_vs_gpu"] = float(rel.max()) if rel.size else None
        line["cpu_baseline"]["port_over_reference_root_time"] = PORT_OVER_REFERENCE
        # the reference binary's throughput on the same cores, estimated from
        # the port's through the measured per-node time ratio (cfg2 root)
        line["cpu_baseline"]["reference_estimate_bounds_per_s"] = round(
            cb["value"] * PORT_OVER_REFERENCE["cfg2 G(100, 0.3) root"], 1)
        progress("CPU baseline done")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
