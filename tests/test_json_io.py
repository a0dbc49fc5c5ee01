"""JSON I/O of TTs (host only, no GPU): json_to_mps / json_to_mpo
(src/tt_tools.jl:765-785) on hand-written fixtures in the layout JSON3.write
gives the reference's structs, and the writers as their inverses."""
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return f.read()


def test_json_to_mps_fixture():
    import ttround as T
    x = T.json_to_mps(_load("ttvector_small.json"))
    assert x.N == 2 and x.ttv_dims == (2, 3) and x.ttv_rks == [1, 2, 1] and x.ttv_ot == [1, 0]
    # flat list read column-major into (n, r, r'): core1[i, 0, b] = v[i + 2b]
    np.testing.assert_array_equal(x.ttv_vec[0][:, 0, :], [[1.0, 3.0], [2.0, 4.0]])
    np.testing.assert_array_equal(x.ttv_vec[1][:, :, 0], [[5.0, 8.0], [6.0, 9.0], [7.0, 10.0]])
    # full tensor T[i1, i2] = sum_a core1[i1, 0, a] core2[i2, a, 0], worked by hand
    full = np.einsum("ia,ja->ij", x.ttv_vec[0][:, 0, :], x.ttv_vec[1][:, :, 0])
    np.testing.assert_array_equal(full, [[29.0, 33.0, 37.0], [42.0, 48.0, 54.0]])


def test_json_to_mpo_fixture():
    import ttround as T
    A = T.json_to_mpo(json.loads(_load("ttoperator_small.json")))
    assert A.N == 2 and A.tto_dims == (2, 2) and A.tto_rks == [1, 2, 1]
    # core1[i, j, 0, b] = v[i + 2 j + 4 b]
    np.testing.assert_array_equal(A.tto_vec[0][:, :, 0, 0], [[1.0, 3.0], [2.0, 4.0]])
    np.testing.assert_array_equal(A.tto_vec[0][:, :, 0, 1], [[0.0, -1.0], [1.0, 2.0]])
    np.testing.assert_array_equal(A.tto_vec[1][:, :, 1, 0], [[1.0, -2.0], [1.5, 3.0]])


def test_json_roundtrip_and_errors():
    import ttround as T
    rng = np.random.default_rng(5)
    dims, rks = (2, 3, 4), [1, 2, 3, 1]
    x = T.TTvector(3, [rng.standard_normal((dims[k], rks[k], rks[k + 1])) for k in range(3)], dims, rks, [1, 1, 0])
    y = T.json_to_mps(json.dumps(T.mps_to_json(x)))
    assert y.ttv_rks == x.ttv_rks and y.ttv_ot == x.ttv_ot
    for a, b in zip(x.ttv_vec, y.ttv_vec):
        np.testing.assert_array_equal(a, b)
    A = T.TToperator(2, [rng.standard_normal((2, 2, 1, 3)), rng.standard_normal((2, 2, 3, 1))], (2, 2), [1, 3, 1])
    B = T.json_to_mpo(T.mpo_to_json(A))
    for a, b in zip(A.tto_vec, B.tto_vec):
        np.testing.assert_array_equal(a, b)
    bad = T.mps_to_json(x)
    bad["ttv_vec"][1] = bad["ttv_vec"][1][:-1]
    with pytest.raises(T.DimensionMismatch):
        T.json_to_mps(bad)
