"""The reference's Python front end (biqbin_base.py:91-313) on this library:
MaxCutSolver / QUBOSolver return the reference's result dictionary, and
QUBOSolver's optimize_input divides the instance by the gcd of its entries
and multiplies the Max-Cut value back.  Input handling runs on the CPU; the
solves are marked gpu and checked against the reference's fixtures."""
import json
import os

import numpy as np
import pytest

from fixture_util import fixture_value
from parallel_biqbin_maxcut_amd import instances as I
from parallel_biqbin_maxcut_amd.solvers import MaxCutSolver, QUBOSolver


def _kcluster(golden, name="kcluster40_025_10_1.json"):
    base = os.path.join(golden, "qubos", name)
    return base, json.load(open(base + "-expected_output.json"))


def test_optimize_input_divides_by_the_gcd(golden):
    """kcluster40's symmetric part has entries -595, 5 and 55: gcd 5, and the
    Max-Cut graph of Q / 5 is the graph of Q divided by 5 (the map is linear)."""
    base, _ = _kcluster(golden)
    plain = QUBOSolver(base, None).read_data()
    s = QUBOSolver(base, None, optimize_input=True)
    scaled = s.read_data()
    assert s.gcd == 5
    assert np.array_equal(scaled * 5, plain)
    assert np.all(scaled == np.round(scaled))


def test_optimize_input_keeps_a_gcd_of_one():
    Q = np.array([[2.0, 3.0], [3.0, -4.0]])
    s = QUBOSolver(Q, None, optimize_input=True)
    assert np.array_equal(s.read_data(), I.qubo_to_maxcut_adj(Q)) and s.gcd == 1


def test_non_integer_qubo_is_rejected():
    with pytest.raises(ValueError, match="integers"):
        QUBOSolver(np.array([[0.5, 0.0], [0.0, 1.0]]), None).read_data()


def test_maxcut_inputs(golden):
    """A rudy path and its adjacency-JSON copy read to the same matrix."""
    path = os.path.join(golden, "rudy", "g05_60.0")
    a = MaxCutSolver(path, None).read_data()
    b = MaxCutSolver(path + ".json", None).read_data()
    assert a.shape == (60, 60) and np.array_equal(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("optimize_input", [False, True])
def test_qubo_solver_kcluster40(golden, optimize_input):
    base, exp = _kcluster(golden)
    r = QUBOSolver(base, os.path.join(golden, "params"), optimize_input=optimize_input, batch=16).run()
    assert r["maxcut"]["computed_val"] == pytest.approx(exp["maxcut"]["computed_val"], abs=1e-6)
    assert r["qubo"]["computed_val"] == pytest.approx(exp["qubo"]["computed_val"], abs=1e-6)
    assert r["qubo"]["cardinality"] == exp["qubo"]["cardinality"]
    assert len(r["maxcut"]["x"]) == len(exp["maxcut"]["x"]) and len(r["qubo"]["x"]) == len(exp["qubo"]["x"])
    assert r["qubo"]["solution"] == [i + 1 for i, v in enumerate(r["qubo"]["x"]) if v]
    assert r["meta_data"]["parameters"]["gcd"] == (5 if optimize_input else 1)


@pytest.mark.gpu
def test_maxcut_solver_g05_60(golden):
    path = os.path.join(golden, "rudy", "g05_60.0")
    r = MaxCutSolver(path, os.path.join(golden, "params"), warm=True).run()
    assert r["maxcut"]["computed_val"] == fixture_value(path + "-expected_output")
    assert r["meta_data"]["instance"] == "g05_60.0" and not r["meta_data"]["stopped"]
