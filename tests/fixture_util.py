"""Small helpers shared by the tests: brute-force Max-Cut and readers of the
reference's expected-output fixtures."""
import numpy as np


def laplacian(adj):
    return np.diag(adj.sum(1)) - adj


def brute_maxcut(adj):
    """Maximum cut by enumeration (the last vertex on side 0)."""
    n = adj.shape[0]
    Lap = laplacian(adj)
    bits = np.arange(1 << (n - 1))[:, None]
    x = ((bits >> np.arange(n - 1)) & 1).astype(float)
    s = 2 * np.hstack([x, np.zeros((x.shape[0], 1))]) - 1
    return float(np.max(0.25 * np.einsum("ki,ij,kj->k", s, Lap, s)))


def fixture_solution(path):
    """1-based vertices on side 1 from a reference '-expected_output' file."""
    for line in open(path):
        if line.startswith("Solution"):
            return [int(t) for t in line.split("(")[1].split(")")[0].split()]


def fixture_value(path):
    for line in open(path):
        if line.startswith("Maximum value"):
            return float(line.split("=")[1])


def fixture_root_bound(path):
    for line in open(path):
        if line.startswith("Root node bound"):
            return float(line.split("=")[1])
