"""Entrywise (Hadamard) products on the device and the NTuple methods that
round or sketch them, against the oracle (FP64: 1e-12 relative, products of
two factors bit-exact).  The oracle's NTuple sketch and rounding work factor by
factor, never forming the product, and so does the device
(``ttr_round_rand_tuple``, ``ttr_sketch_tuple`` in both directions);
``ttr_hadamard`` forms the product core by core, and rounding that product with
the same draws must give the factored result."""
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


def test_hadamard_vs_oracle(T):
    say("case: x * y (kron order) and a 3-factor product (tuple order)")
    dims = (3, 4, 2, 3, 2)
    x = rand_tt(dims, [1, 3, 7, 5, 2, 1], 300)
    y = rand_tt(dims, [1, 2, 6, 4, 2, 1], 301)
    z = rand_tt(dims, [1, 3, 3, 3, 2, 1], 302)
    p = to_oracle(to_api(x) * to_api(y))
    ref = O.hadamard([x, y], first_fastest=False)
    assert p.rks == ref.rks and p.ot == [0] * 5
    for k in range(5):
        np.testing.assert_array_equal(p.cores[k], ref.cores[k])
    dev = [T.DeviceTT.upload(to_api(t)) for t in (x, y, z)]
    d3 = T.hadamard_device(dev, first_fastest=True)
    ref3 = O.hadamard([x, y, z], first_fastest=True)
    assert d3.rks == ref3.rks
    for k in range(5):
        np.testing.assert_allclose(d3.core(k), ref3.cores[k], rtol=1e-15, atol=0)
    for d in dev + [d3]:
        d.free()


@pytest.mark.parametrize("orthogonal", [True, False])
def test_ttrand_rounding_tuple_vs_oracle(T, orthogonal):
    say(f"case: ttrand_rounding((x, y), rks) orthogonal={orthogonal}")
    dims = (2, 3, 2, 3, 2, 2)
    x = rand_tt(dims, [1, 2, 4, 4, 3, 2, 1], 310)
    y = rand_tt(dims, [1, 2, 3, 4, 4, 2, 1], 311)
    rks = [1, 3, 8, 10, 8, 3, 1]  # truncating
    out = to_oracle(T.ttrand_rounding((to_api(x), to_api(y)), rks, orthogonal=orthogonal, seed=17))
    prov = device_blocks_provider(dims, rks, orthogonal=orthogonal, seed=17)
    ref = O.ttrand_rounding_tuple([x, y], rks, prov, orthogonal)
    assert out.rks == ref.rks and out.ot == ref.ot
    err = O.rel_err(ref, out)
    say(f"  rel err vs factored oracle: {err:.3e}")
    assert err < 1e-12
    # ranks that cover the product: exact recovery of x .* y
    full = [1] + [a * b for a, b in zip(x.rks[1:-1], y.rks[1:-1])] + [1]
    ex = to_oracle(T.ttrand_rounding((to_api(x), to_api(y)), full, orthogonal=orthogonal, seed=18))
    assert O.rel_err(O.hadamard([x, y]), ex) < 1e-11
    # a one-element tuple is the TT itself
    one = to_oracle(T.ttrand_rounding((to_api(x),), rks, seed=17, orthogonal=orthogonal))
    ref1 = to_oracle(T.ttrand_rounding(to_api(x), rks, seed=17, orthogonal=orthogonal))
    assert one.rks == ref1.rks
    for a, b in zip(one.cores, ref1.cores):
        np.testing.assert_array_equal(a, b)


def test_sketches_of_tuple_vs_oracle(T):
    say("case: tt_recursive_sketch / tt_sketch / tt_combined_sketch of (x, y, z)")
    dims = (2, 3, 2, 2, 3)
    N = len(dims)
    xs = [rand_tt(dims, [1, 2, 3, 3, 2, 1], 320 + i) for i in range(3)]
    api = tuple(to_api(t) for t in xs)
    rks = [1, 6, 9, 9, 5, 1]
    W, srk = T.tt_recursive_sketch(api, rks, seed=5)
    prov = device_blocks_provider(dims, rks, orthogonal=True, seed=5)
    Wr, sr = O.tt_recursive_sketch_tuple(xs, rks, prov, True)
    assert srk == sr
    for k in range(N + 1):
        assert W[k].shape == Wr[k].shape == tuple(t.rks[k] for t in xs) + (sr[k],)
        np.testing.assert_allclose(W[k], Wr[k], rtol=1e-12, atol=1e-12 * np.abs(Wr[k]).max())
    s = 11
    S, sa = T.tt_sketch(api, s, seed=9)
    bv, _ = O.sketch_structure(dims, [1] * (N + 1))
    p = -(-s // bv[0])
    prov = device_blocks_provider(dims, [1] * (N + 1), orthogonal=True, seed=9, p=[p] * N + [1])
    Sr, sb = O.tt_sketch_tuple(xs, s, prov, True)
    assert sa == sb and S.shape == (sa,)
    np.testing.assert_allclose(S, Sr, rtol=1e-12, atol=1e-12 * np.abs(Sr).max())
    sk, Wc, sc = T.tt_combined_sketch(api, rks, 30, seed=5)
    assert len(sk) >= 30 and sc[0] == 1 and Wc[0].shape == (1, 1, 1, 1)
    for k in range(1, N + 1):
        assert Wc[k].shape == tuple(t.rks[k] for t in xs) + (sc[k],)


@pytest.mark.parametrize("orthogonal,three,p", [(True, False, None), (False, True, None), (True, True, [1, 6, 5, 3, 2])])
def test_forward_sketch_of_tuple(T, orthogonal, three, p):
    """tt_recursive_sketch((x, y[, z]), rks; reverse=false): factored on the device
    (the product never formed), against the oracle's factored forward sketch and
    the forward sketch of the explicit product in tuple order, rows reshaped to
    (r_xk, r_yk[, r_zk], s_k)."""
    say(f"case: forward sketch of a product of {3 if three else 2} factors, orthogonal={orthogonal}, p={p}")
    dims = (2, 3, 2, 3)
    fs = [rand_tt(dims, [1, 2, 3, 2, 1], 340), rand_tt(dims, [1, 2, 2, 3, 1], 341)]
    if three:
        fs.append(rand_tt(dims, [1, 2, 3, 3, 1], 342))
    rks = [1, 4, 6, 5, 1]
    W, s = T.tt_recursive_sketch(tuple(to_api(t) for t in fs), rks, reverse=False, seed=12, orthogonal=orthogonal,
                                 p=p)
    prov = device_blocks_provider(dims, rks, orthogonal=orthogonal, seed=12, forward=True, p=p)
    Wf, sf = O.tt_recursive_sketch_tuple_forward(fs, rks, prov, orthogonal, p=p)
    Wr, sr = O.tt_recursive_sketch_forward(O.hadamard(fs, first_fastest=True), rks, prov, orthogonal, p=p)
    assert s == sr == sf
    worst = 0.0
    for k in range(len(dims) + 1):
        ref = Wr[k].reshape(tuple(t.rks[k] for t in fs) + (-1,), order="F")
        assert W[k].shape == ref.shape == Wf[k].shape
        np.testing.assert_allclose(W[k], ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
        worst = max(worst, np.linalg.norm(W[k] - Wf[k]) / np.linalg.norm(Wf[k]))
    say(f"  sketch ranks {s}, worst rel err vs the factored oracle {worst:.3e}")
    assert worst < 1e-12


def test_hadamard_errors(T):
    say("case: ttr_hadamard argument errors")
    x = rand_tt((2, 3, 2), [1, 2, 2, 1], 330)
    y = rand_tt((2, 2, 2), [1, 2, 2, 1], 331)
    with pytest.raises(T.DimensionMismatch):
        to_api(x) * to_api(y)
    with pytest.raises(T.DimensionMismatch):
        T.ttrand_rounding((to_api(x), to_api(y)), 4)
    with pytest.raises(TypeError):
        T.ttrand_rounding((to_api(x), to_api(x)))
    dx = T.DeviceTT.upload(to_api(x))
    with pytest.raises(Exception):
        T.hadamard_device([dx] * 9)
    dx.free()


def _formed_round(T, xs, rks, **kw):
    dev = [T.DeviceTT.upload(to_api(t)) for t in xs]
    prod = T.hadamard_device(dev, first_fastest=True)
    out = T.round_rand_batch([([1.0], [prod], rks)], **kw)[0]
    res = to_oracle(out.to_host())
    for d in dev + [prod, out]:
        d.free()
    return res


def _factored_round(T, items, **kw):
    devs = [[T.DeviceTT.upload(to_api(t)) for t in xs] for xs, _ in items]
    outs = T.round_rand_tuple_batch([(d, r) for d, (_, r) in zip(devs, items)], **kw)
    res = [to_oracle(o.to_host()) for o in outs]
    for d in [x for ds in devs for x in ds] + outs:
        d.free()
    return res


def test_factored_equals_formed_product(T):
    """Two items in one batch (2 and 3 factors): the factored rounding equals the
    rounding of the formed product on the same draws, and the factored oracle."""
    dims = (3, 4, 3, 4, 3)
    a = [rand_tt(dims, [1, 3, 9, 12, 3, 1], 360), rand_tt(dims, [1, 3, 8, 7, 3, 1], 361)]
    b = [rand_tt(dims, [1, 2, 4, 5, 3, 1], 362 + j) for j in range(3)]
    ra, rb = [1, 3, 12, 24, 9, 1], [1, 3, 10, 20, 8, 1]
    for prec, acc in (("f64", "sketch"), ("f32", "f64")):
        say(f"case: factored vs formed product, sketch {prec}, accumulation {acc}")
        got = _factored_round(T, [(a, ra), (b, rb)], orthogonal=False, seed=21, sketch_prec=prec, sketch_accum=acc)
        for xs, r, g in ((a, ra, got[0]), (b, rb, got[1])):
            formed = _formed_round(T, xs, r, orthogonal=False, seed=21, sketch_prec=prec, sketch_accum=acc)
            blocks = {k: T.sketch_blocks(dims, r, k, orthogonal=False, seed=21, sketch_prec=prec) for k in range(5)}
            ref = O.ttrand_rounding_tuple(xs, r, blocks.__getitem__, False)
            e1, e2 = O.rel_err(formed, g), O.rel_err(ref, g)
            say(f"  {len(xs)} factors: vs formed product {e1:.3e}, vs factored oracle {e2:.3e}")
            assert g.rks == formed.rks == ref.rks
            assert e1 < 1e-12 and e2 < 1e-12


def _entry(tt, idx):
    v = np.ones((1, 1))
    for k, i in enumerate(idx):
        v = v @ tt.cores[k][i]
    return float(v[0, 0])


def test_product_too_large_to_form(T):
    """Three factors with representation ranks 100 (each 100 scaled copies of one
    rank-1 TT): the product has ranks 10^6, its cores far beyond 2^31 entries,
    so ttr_hadamard refuses it; the factored rounding to ranks 2 recovers the
    (rank-1) product, checked entry by entry at random multi-indices."""
    say("case: NTuple rounding of a product with ranks 10^6")
    dims = (8,) * 6
    rng = np.random.default_rng(370)
    facs = []
    for f in range(3):
        y = O.rand_tt(dims, 1, rng)
        c = rng.uniform(0.5, 1.5, 100) / 100.0
        x = O.tt_lincomb([y] * 100, list(c))
        assert max(x.rks) == 100
        facs.append(x)
    dev = [T.DeviceTT.upload(to_api(t)) for t in facs]
    with pytest.raises(ValueError):
        T.hadamard_device(dev, first_fastest=True)
    for d in dev:
        d.free()
    rks = [1] + [2] * 5 + [1]
    out = _factored_round(T, [(facs, rks)], orthogonal=True, seed=4)[0]
    bv, p = O.sketch_structure(dims, rks, None)
    assert out.rks == [1] + [min(8 ** (k + 1), bv[k + 1] * p[k + 1]) for k in range(5)] + [1]
    worst = 0.0
    for _ in range(200):
        idx = tuple(int(i) for i in rng.integers(0, 8, 6))
        ref = np.prod([_entry(x, idx) for x in facs])
        worst = max(worst, abs(_entry(out, idx) - ref) / max(abs(ref), 1e-300))
    say(f"  max relative entry error over 200 samples: {worst:.3e}")
    assert worst < 1e-10
