"""Instance formats and seeded generators.

* rudy edge lists as read by the reference (src/process_input.c:201-253):
  first line "n m", then m lines "i j w" (1-based, integer weights).
* adjacency / QUBO JSON as read by the reference's Python front end
  (biqbin_base.py:33-45, 186-219): {"qubo"|"adjacency": {"shape", "row", "col", "data"}}.
* QUBO -> Max-Cut on n+1 vertices and back (biqbin_base.py:230-281).
* Seeded SplitMix64 generators for the benchmark configurations.
"""
import json

import numpy as np

MASK64 = (1 << 64) - 1


def read_rudy(path):
    """Adjacency matrix of a rudy file; non-integer weights raise ValueError
    (process_input.c:243-244)."""
    with open(path) as f:
        toks = f.read().split()
    nv, ne = int(toks[0]), int(toks[1])
    if nv <= 0:
        raise ValueError("Number of vertices has to be positive")
    adj = np.zeros((nv, nv))
    p = 2
    for _ in range(ne):
        i, j, w = int(toks[p]), int(toks[p + 1]), float(toks[p + 2])
        p += 3
        if not (1 <= i <= nv and 1 <= j <= nv):
            raise ValueError("Problem with edge. Vertex not in range")
        if float(int(w)) != w:
            raise ValueError("Edge weight value error! All edge weights need to be integers!")
        adj[i - 1, j - 1] = w
        adj[j - 1, i - 1] = w
    return adj


def write_rudy(path, adj):
    adj = np.asarray(adj)
    n = adj.shape[0]
    edges = [(i, j, adj[i, j]) for i in range(n) for j in range(i + 1, n) if adj[i, j] != 0]
    with open(path, "w") as f:
        f.write(f"{n} {len(edges)}\n")
        for i, j, w in edges:
            f.write(f"{i + 1} {j + 1} {int(w)}\n")


def read_sparse_json(path, key):
    with open(path) as f:
        d = json.load(f)[key]
    m = np.zeros(tuple(d["shape"]))
    for r, c, v in zip(d["row"], d["col"], d["data"]):
        m[r, c] += v
    return m


def qubo_to_maxcut_adj(Q):
    """biqbin_base.py:230-258: symmetric part, off-diagonal -> edges, and a
    homogenising vertex joined with weight -rowsum."""
    q_sym = 0.5 * (Q.T + Q)
    if not np.all(q_sym == q_sym.astype(np.int64)):
        raise ValueError("All QUBO values need to be integers!")
    qe = -q_sym.sum(axis=1)
    a = q_sym.copy()
    np.fill_diagonal(a, 0)
    n = Q.shape[0]
    out = np.zeros((n + 1, n + 1))
    out[:n, :n] = a
    out[:n, n] = qe
    out[n, :n] = qe
    return out


def maxcut_solution_to_qubo(sol_vertices, n):
    """biqbin_base.py:260-281: 1-based cut vertices of the n+1 graph -> QUBO x."""
    idx = np.array(sol_vertices, dtype=int) - 1
    x = np.ones(n + 1)
    x[idx] = -1
    x *= -x[-1]
    y = 0.5 * (x + 1)[:-1]
    return y.astype(int)


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014) stream."""

    def __init__(self, seed):
        self.s = seed & MASK64

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & MASK64
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next() >> 11) * (1.0 / (1 << 53))

    def randint(self, lo, hi):
        """Integer uniform in [lo, hi]."""
        return lo + (self.next() >> 32) % (hi - lo + 1)


def gnp_adj(n, p, seed, weights="unit"):
    """G(n, p) adjacency; weights: 'unit', 'pm1' ({-1,+1}), 'int10' ({-10..10}\\{0}),
    'zpm1' ({-1,0,+1}).  Edges are drawn for i<j in row order."""
    r = SplitMix64(seed)
    adj = np.zeros((n, n))
    for i in range(n):
        for j in range(i + 1, n):
            if r.uniform() < p:
                if weights == "unit":
                    w = 1
                elif weights == "pm1":
                    w = 1 if r.next() >> 63 else -1
                elif weights == "int10":
                    w = r.randint(1, 10) * (1 if r.next() >> 63 else -1)
                elif weights == "zpm1":
                    w = r.randint(-1, 1)
                else:
                    raise ValueError(weights)
                adj[i, j] = adj[j, i] = w
    return adj


def random_qubo(n, density, seed, lo=-100, hi=100):
    """Symmetric integer Q: `density` of the upper triangle (diagonal included)
    uniform in [lo, hi], mirrored."""
    r = SplitMix64(seed)
    Q = np.zeros((n, n))
    for i in range(n):
        for j in range(i, n):
            if r.uniform() < density:
                Q[i, j] = Q[j, i] = r.randint(lo, hi)
    return Q


def random_fixings(n0, nfix_max, count, seed):
    """`count` B&B nodes: each fixes f ~ U{0..nfix_max} distinct vertices among
    the first n0-1 to random 0/1 values.  Returns (xfixed, solX) int32 [count, n0-1]."""
    r = SplitMix64(seed)
    B = n0 - 1
    xf = np.zeros((count, B), dtype=np.int32)
    sx = np.zeros((count, B), dtype=np.int32)
    for c in range(count):
        f = r.randint(0, nfix_max)
        perm = list(range(B))
        for i in range(B - 1, 0, -1):
            j = r.randint(0, i)
            perm[i], perm[j] = perm[j], perm[i]
        for v in perm[:f]:
            xf[c, v] = 1
            sx[c, v] = r.next() >> 63
    return xf, sx
