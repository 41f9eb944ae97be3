"""The CPU FP64 restatement of the cut model (oracle/orc_mlp.py) -- the
checker of the HIP training kernels -- is pinned here by central finite
differences of its own loss (independent of the analytic backward pass it
restates), on every parameter block."""
import numpy as np

import orc_mlp as M


def test_mlp_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    npar = sum(int(np.prod(s)) for _, s in M.SIZES)
    p = rng.normal(size=npar) * 0.08
    d = M.unpack(p)
    d["W4"][2:] = 0.0
    d["b4"][2:] = 0.0
    p = M.pack(d)
    x = rng.normal(size=(8, M.NF))
    x[:, 16:] = 0.0
    t = np.stack([np.abs(rng.normal(size=8)), (rng.random(8) < 0.5).astype(float)], axis=1)
    _, g = M.loss_and_grad(p, x, t)
    offs, o = {}, 0
    for name, shp in M.SIZES:
        offs[name] = (o, int(np.prod(shp)))
        o += int(np.prod(shp))
    checked = 0
    for name, (o, k) in offs.items():
        idx = rng.choice(k, size=min(k, 40), replace=False) + o
        if name in ("W4", "b4"):
            idx = [i for i in idx if (i - o) // (M.H if name == "W4" else 1) < 2] + ([o, o + 1] if name == "b4" else [o, o + 5])
        for i in idx:
            h = 1e-6
            pp, pm = p.copy(), p.copy()
            pp[i] += h
            pm[i] -= h
            fd = (M.loss_only(pp, x, t) - M.loss_only(pm, x, t)) / (2 * h)
            assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-8, (name, i, fd, g[i])
            checked += 1
    assert checked >= 240


def test_mlp_oracle_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -3.14159, 0.0])
    r = M.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -7 and r[4] == 0.0
    assert abs(r[3] + 3.140625) < 1e-12
