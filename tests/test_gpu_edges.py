"""Edge cases of the rounding entry points against the oracle: zero TTs (the
reference returns zeros_tt with the input ranks, src/tt_rounding.jl:178-180,
222-224), a single core, two cores, unit mode sizes and rank-one TTs."""
import sys

import numpy as np
import pytest

import oracle as O
from helpers import device_blocks_provider, to_api, to_oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def T():
    import ttround
    ttround.default_handle()
    return ttround


def rand_tt(dims, rks, seed):
    return O.rand_tt(dims, rks, np.random.default_rng(seed))


@pytest.mark.parametrize("direction", ["left", "right"])
def test_zero_tt(T, direction):
    say(f"case: zero TT rounding direction={direction}")
    z = O.zeros_tt((3, 2, 4, 2), [1, 3, 4, 2, 1])
    r = to_oracle(T.tt_rounding(to_api(z), tol=1e-8, direction=direction))
    ref = O.tt_rounding(z, tol=1e-8, direction=direction)
    assert r.rks == ref.rks == z.rks
    for c in r.cores:
        assert not np.any(c)
    assert T.norm(to_api(z)) == 0.0
    x = rand_tt((3, 2, 4, 2), [1, 3, 4, 2, 1], 500)
    assert T.dot(to_api(z), to_api(x)) == 0.0


def test_single_and_two_cores(T):
    say("case: d = 1 and d = 2")
    x1 = rand_tt((7,), [1, 1], 501)
    r1 = to_oracle(T.tt_rounding(to_api(x1), tol=1e-10))
    assert r1.rks == [1, 1]
    np.testing.assert_allclose(O.ttv_to_tensor(r1), O.ttv_to_tensor(x1), rtol=1e-14, atol=1e-15)
    x2 = O.tt_add(rand_tt((5, 6), [1, 3, 1], 502), rand_tt((5, 6), [1, 2, 1], 503))
    r2 = to_oracle(T.tt_rounding(to_api(x2), tol=1e-10))
    ref2 = O.tt_rounding(x2, tol=1e-10)
    assert r2.rks == ref2.rks == [1, 5, 1]
    assert O.rel_err(ref2, r2) < 1e-12
    rks = [1, 5, 1]
    y2 = to_oracle(T.ttrand_rounding(to_api(x2), rks, seed=3))
    prov = device_blocks_provider((5, 6), rks, orthogonal=True, seed=3)
    refy = O.ttrand_rounding(x2, rks, prov, True)
    assert y2.rks == refy.rks
    assert O.rel_err(refy, y2) < 1e-12
    assert abs(T.norm(to_api(x2)) - O.tt_norm(x2)) <= 1e-12 * O.tt_norm(x2)


def test_unit_modes_and_rank_one(T):
    say("case: unit mode sizes, rank-one TT")
    dims = (2, 1, 3, 1, 2)
    x = O.tt_add(rand_tt(dims, [1, 2, 2, 2, 2, 1], 504), rand_tt(dims, [1, 2, 2, 2, 2, 1], 505))
    r = to_oracle(T.tt_rounding(to_api(x), tol=1e-12))
    ref = O.tt_rounding(x, tol=1e-12)
    assert r.rks == ref.rks
    assert O.rel_err(ref, r) < 1e-12
    one = rand_tt((3, 3, 3, 3), [1, 1, 1, 1, 1], 506)
    y = to_oracle(T.ttrand_rounding(to_api(one), [1, 2, 2, 2, 1], seed=4))
    assert O.rel_err(one, y) < 1e-13
    s = to_oracle(T.tt_rounding(to_api(O.tt_add(one, one)), tol=1e-12))
    assert s.rks == [1, 1, 1, 1, 1]
    assert O.rel_err(O.tt_scale(2.0, one), s) < 1e-13


def test_trim_keeps_live_objects(T):
    """ttr_trim returns the idle workspace and freed storage; live TTs stay valid
    and a call after it gives the same bits as before it."""
    say("case: ttr_trim between two batched SVD roundings")
    xs = [T.DeviceTT.randn((4,) * 6, [1, 4, 16, 12, 16, 4, 1], 900 + i) for i in range(3)]
    s = T.lincomb_device([1.0, 0.5, 0.25], xs)
    r1 = T.round_svd_batch([s, s], tol=1e-10)
    c1 = [r1[0].core(k) for k in range(6)]
    for o in r1:
        o.free()
    T.default_handle().trim()
    r2 = T.round_svd_batch([s, s], tol=1e-10)
    for k in range(6):
        np.testing.assert_array_equal(r2[0].core(k), c1[k])
        np.testing.assert_array_equal(r2[1].core(k), c1[k])
    assert xs[0].core(2).shape == (4, 16, 12)
    for o in r2 + xs + [s]:
        o.free()
