"""The serve bench's error summary: distances computed as solves finish
(against the reference of the request's distinct sequence) give the same
figures as distances computed afterwards from stored velocities."""
import numpy as np

import bench


def test_serve_errors_agree_between_forms():
    rng = np.random.default_rng(3)
    ref = {(k, j): (rng.normal(size=50), 10 + k) for k in range(3) for j in range(2)}
    key = [0, 1, 2, 1]   # request 3 repeats sequence 1
    stored, streamed = {}, {}
    for k, s in enumerate(key):
        for j in range(2):
            v = ref[(s, j)][0] * (1 + 0.7e-6 * (k + 1) * (j + 1))
            stored[(k, j)] = (v, 10 + s, 0)
            streamed[(k, j)] = (bench.rel_l2(v, ref[(s, j)][0]), 10 + s, 0, ref[(s, j)][1])
    ref_by_request = {(k, j): ref[(key[k], j)] for k in range(len(key)) for j in range(2)}
    a = bench.serve_errors(stored, ref_by_request)
    b = bench.serve_errors(streamed, None)
    assert a == b
    assert a["solves"] == 8
    assert a["worst_solve"]["request"] == 3 and a["worst_solve"]["solve"] == 1
    assert np.isclose(a["max_rel_l2"], 5.6e-6)
    assert a["solves_above_1e-6"] == 7
