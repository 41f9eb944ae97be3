"""CPU FP64 restatement of the cut model's forward and backward pass
(TEST INFRASTRUCTURE ONLY: the checker of the library's HIP training
kernels; nothing in the product path uses it).

Model (parallel_biqbin_maxcut_amd/csrc/bqb_learn.h): x [32] ->
h1 = relu(W1 x + b1) -> h2 = relu(W2 h1 + b2) -> h3 = relu(W3 h2 + b3) ->
o = W4 h3 + b4 (rows 0, 1 used).  Loss per row: (o0 - t0)^2 +
BCE(sigmoid(o1), t1), averaged over the rows.  Flat parameter layout
W1[256][32] b1 W2[256][256] b2 W3[256][256] b3 W4[16][256] b4[16].

bf16=True rounds what the device keeps in bf16 (the features, the weights
the matrix cores read, the activations and the gradients passed between
layers) so that the comparison measures the FP32 accumulation only."""
import numpy as np

H, NF, NO = 256, 32, 16
SIZES = [("W1", (H, NF)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, H)), ("b3", (H,)),
         ("W4", (NO, H)), ("b4", (NO,))]


def unpack(p):
    out, o = {}, 0
    for name, shp in SIZES:
        k = int(np.prod(shp))
        out[name] = np.asarray(p[o:o + k], dtype=np.float64).reshape(shp)
        o += k
    return out


def pack(d):
    return np.concatenate([d[name].reshape(-1) for name, _ in SIZES])


def bf16(a):
    """Round to bfloat16 (nearest even), returned as float64."""
    a32 = np.asarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(p, x, bf=False):
    w = unpack(p)
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    W = {k: r(v) if k.startswith("W") else v for k, v in w.items()}
    x = r(x)
    a1 = x @ W["W1"].T + w["b1"]
    h1 = r(np.maximum(a1, 0))
    a2 = h1 @ W["W2"].T + w["b2"]
    h2 = r(np.maximum(a2, 0))
    a3 = h2 @ W["W3"].T + w["b3"]
    h3 = r(np.maximum(a3, 0))
    o = h3 @ W["W4"].T + w["b4"]
    return dict(x=x, h1=h1, h2=h2, h3=h3, o=o, W=W)


def loss_and_grad(p, x, t, bf=False):
    """Mean loss and its gradient with respect to the flat parameters."""
    f = forward(p, x, bf)
    r = bf16 if bf else (lambda a: a)
    M = x.shape[0]
    o0, o1 = f["o"][:, 0], f["o"][:, 1]
    s = 1.0 / (1.0 + np.exp(-o1))
    loss = np.mean((o0 - t[:, 0]) ** 2 + np.maximum(o1, 0) - o1 * t[:, 1] + np.log1p(np.exp(-np.abs(o1))))
    d4 = np.zeros((M, NO))
    d4[:, 0] = 2.0 * (o0 - t[:, 0]) / M
    d4[:, 1] = (s - t[:, 1]) / M
    d4 = r(d4)
    W = f["W"]
    d3 = r(d4 @ W["W4"]) * (f["h3"] > 0)
    d2 = r(d3 @ W["W3"]) * (f["h2"] > 0)
    d1 = r(d2 @ W["W2"]) * (f["h1"] > 0)
    g = dict(W1=d1.T @ f["x"], b1=d1.sum(0), W2=d2.T @ f["h1"], b2=d2.sum(0), W3=d3.T @ f["h2"], b3=d3.sum(0),
             W4=d4.T @ f["h3"], b4=d4.sum(0))
    g["W4"][2:] = 0.0
    g["b4"][2:] = 0.0
    return loss, pack(g)


def loss_only(p, x, t):
    f = forward(p, x)
    o0, o1 = f["o"][:, 0], f["o"][:, 1]
    return np.mean((o0 - t[:, 0]) ** 2 + np.maximum(o1, 0) - o1 * t[:, 1] + np.log1p(np.exp(-np.abs(o1))))
