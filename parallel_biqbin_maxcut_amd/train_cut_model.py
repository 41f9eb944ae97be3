"""Trains the cut model on solver-generated data and measures its gain.

    python -m parallel_biqbin_maxcut_amd.train_cut_model OUT.npz [graphs] [epochs]

1. Data: the root nodes of `graphs` (default 2000) G(80, p ~ U[0.1, 0.9])
   with weights in {-1, 0, +1}, bounded on the GPU with the reference's
   parameters while the context records the features of the new cuts of
   every separation; labels from each node's final cut set (learn.py).
2. Training: Adam on the GPU with the library's bf16 forward/backward
   kernels (bqb_mlp_train), 90 % of the graphs' rows, the rest held out.
3. Gain: 100 held-out G(120, 0.5) +-1 graphs, root bound with the plain
   cold-start bundle and with the model, 6 outer rounds each.
Writes the parameters (float32, BQB_MLP_NPARAM) and a summary to OUT.npz and
prints the summary as one JSON line."""
import json
import os
import sys
import time

if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

import numpy as np

from . import _native as N
from . import api, learn


def main(argv):
    out = argv[1] if len(argv) > 1 else "cut_model.npz"
    ngraph = int(argv[2]) if len(argv) > 2 else 2000
    epochs = int(argv[3]) if len(argv) > 3 else 6
    N.check(N.lib().bqb_init(0))
    P = api.make_params()
    P.TriIneq = 800                                   # adjust_TriIneq: 10 * n0
    t0 = time.perf_counter()
    graphs = learn.training_graphs(ngraph, 7)
    feat, target = learn.generate(graphs, P, seed=1, workers=16,
                                  progress=lambda d, t: print(f"data: {d}/{t} graphs", flush=True))
    t_data = time.perf_counter() - t0
    print(f"data: {len(feat)} rows from {ngraph} graphs in {t_data:.1f} s, active {target[:, 1].mean():.3f}",
          flush=True)
    rng = np.random.default_rng(5)
    perm = rng.permutation(len(feat))
    cut = int(0.9 * len(feat))
    tr, va = perm[:cut], perm[cut:]
    mlp = learn.MLP(seed=3)
    t0 = time.perf_counter()
    losses = mlp.train(feat[tr], target[tr], epochs=epochs, batch=2048, lr=1e-3, seed=9)
    t_train = time.perf_counter() - t0
    _, val_loss = mlp.gradient(feat[va][:200000], target[va][:200000])
    pred = mlp.forward(feat[va][:200000])
    acc = float(np.mean((pred[:, 1] > 0) == (target[va][:200000, 1] > 0.5)))
    print(f"train: {epochs} epochs in {t_train:.1f} s, loss {losses.round(4).tolist()}, held-out loss "
          f"{val_loss:.4f}, activity accuracy {acc:.3f}", flush=True)
    Pe = api.make_params(max_outer_iter=6, extra_iter=0)
    Pe.TriIneq = 1200
    eg = learn.eval_graphs(100)
    gains = {}
    for drop in (0.0, 0.05):
        t0 = time.perf_counter()
        _, s = learn.evaluate_gain(mlp, eg, Pe, drop=drop, workers=16)
        s["seconds"] = round(time.perf_counter() - t0, 1)
        gains[f"drop_{drop}"] = s
        print(f"gain drop {drop}: {json.dumps(s)}", flush=True)
    summary = {"rows": int(len(feat)), "graphs": ngraph, "data_s": round(t_data, 1), "train_s": round(t_train, 1),
               "epochs": epochs, "loss": losses.tolist(), "heldout_loss": val_loss, "activity_accuracy": acc,
               "gain": gains}
    np.savez(out, params=mlp.params(), summary=json.dumps(summary))
    mlp.close()
    print(json.dumps(summary), flush=True)


if __name__ == "__main__":
    main(sys.argv)
