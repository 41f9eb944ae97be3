"""Python front end with the reference's interface (biqbin_base.py:91-313):
MaxCutSolver and QUBOSolver, on this library's single-process node queue
(bqb_solve_opt) instead of the MPI run.

Inputs are arrays or files: a Max-Cut instance as an adjacency matrix, a
rudy file or an adjacency JSON (biqbin_base.py:48-89); a QUBO as a matrix or
a JSON file with a "qubo" entry (biqbin_base.py:186-220).  `params` is the
path of a parameter file (src/process_input.c:153-180) or a
BiqBinParameters.  Further keyword arguments select the solver form
(api.solve: batch, warm, ipm_warm, slices, certify, model, graphs).

run() returns the reference's result dictionary: "time", "maxcut" with
"computed_val" and "solution" (1-based vertices on side 1; the last vertex is
on side 0), for QUBOs also "maxcut"/"x" and "qubo" with "computed_val",
"solution", "x" and "cardinality", and "meta_data"."""
import math
import os
import time
from functools import reduce

import numpy as np

from . import api, instances

__all__ = ["MaxCutSolver", "QUBOSolver"]


def _params(params, nv, time_limit):
    if isinstance(params, (str, os.PathLike)):
        p = api.read_params(params)
    else:
        p = type(params).from_buffer_copy(params)   # the caller's struct stays as given
    if p.adjust_TriIneq:
        p.TriIneq = 10 * nv                       # src/process_input.c:135-136
    if time_limit:
        p.time_limit = int(time_limit)
    return p


class MaxCutSolver:
    """biqbin_base.py:91-184 on one GPU."""
    solver_name = "biqbin-mi355x MaxCut"

    def __init__(self, instance, params, time_limit=0, **solve_options):
        self.instance = instance
        self.params = params
        self.time_limit = time_limit
        self.solve_options = dict(solve_options)

    def problem_instance_name(self):
        return os.path.basename(self.instance) if isinstance(self.instance, (str, os.PathLike)) else "array"

    def read_data(self):
        """The Max-Cut adjacency matrix (biqbin_base.py:110-116)."""
        a = self.instance
        if isinstance(a, (str, os.PathLike)):
            return instances.read_sparse_json(a, "adjacency") if str(a).endswith(".json") else api.read_rudy(a)
        return np.asarray(a, dtype=np.float64)

    def _solve(self, adj):
        SP = api.maxcut_L(adj)
        P = _params(self.params, SP.shape[0], self.time_limit)
        t0 = time.perf_counter()
        st, sol = api.solve(SP, P, **self.solve_options)
        if st.status:
            raise RuntimeError(f"node bound failed with status {st.status}")
        return st, sol, time.perf_counter() - t0

    def _result(self, st, sol, secs):
        return {"time": secs,
                "maxcut": {"computed_val": float(st.best),
                           "solution": [i + 1 for i in range(len(sol)) if sol[i] == 1]},
                "meta_data": {"solver": self.solver_name, "instance": self.problem_instance_name(),
                              "nodes": int(st.nodes), "root_bound": float(st.root_bound),
                              "stopped": bool(st.stopped),
                              "parameters": {"time_limit": self.time_limit if self.time_limit > 0 else None}}}

    def run(self):
        """Solves to the proven optimum; the result dictionary of
        biqbin_base.py:133-146."""
        st, sol, secs = self._solve(self.read_data())
        return self._result(st, sol, secs)


class QUBOSolver(MaxCutSolver):
    """biqbin_base.py:222-313: min x'Qx over x in {0,1}^n through Max-Cut on
    n + 1 vertices."""
    solver_name = "biqbin-mi355x QUBO"

    def __init__(self, instance, params, optimize_input=False, time_limit=0, **solve_options):
        super().__init__(instance, params, time_limit, **solve_options)
        self.optimize_input = optimize_input
        self.gcd = 1

    def qubo(self):
        q = self.instance
        if isinstance(q, (str, os.PathLike)):
            return instances.read_sparse_json(q, "qubo")
        return np.asarray(q, dtype=np.float64)

    def read_data(self):
        """The Max-Cut graph of the QUBO (biqbin_base.py:230-258); with
        optimize_input the integer symmetric part is first divided by the gcd
        of its entries."""
        Q = self.qubo()
        q_sym = 0.5 * (Q.T + Q)
        if not np.all(q_sym == q_sym.astype(np.int64)):
            raise ValueError("All QUBO values need to be integers!")
        self.gcd = 1
        if self.optimize_input:
            g = reduce(math.gcd, {int(v) for v in q_sym.flatten()})
            if g > 1:
                self.gcd = g
                return instances.qubo_to_maxcut_adj(q_sym / g)
        return instances.qubo_to_maxcut_adj(Q)

    def run(self):
        st, sol, secs = self._solve(self.read_data())
        result = self._result(st, sol, secs)
        Q = self.qubo()
        n = Q.shape[0]
        cut = result["maxcut"]["solution"]
        y = instances.maxcut_solution_to_qubo(cut, n)              # biqbin_base.py:260-281
        xx = np.zeros(n + 1, dtype=int)
        xx[np.array(cut, dtype=int) - 1] = 1
        result["maxcut"]["computed_val"] *= self.gcd
        result["maxcut"]["x"] = xx.tolist()
        result["qubo"] = {"computed_val": float(y @ Q @ y),
                          "solution": (np.nonzero(y)[0] + 1).tolist(),
                          "x": y.astype(int).tolist(),
                          "cardinality": float(y.sum())}
        result["meta_data"]["parameters"]["optimize_input"] = self.optimize_input
        result["meta_data"]["parameters"]["gcd"] = self.gcd
        return result
