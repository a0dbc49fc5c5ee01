"""Parity of the HIP path (through the C-ABI) with the FP64 oracle.

Tolerances: FP64 paths within 1e-12 relative (represented tensor, Frobenius
norm); FP32 sketches 1e-6; BF16 sketches 1e-6 when the sketch rank covers the
true rank (exact recovery), otherwise against the oracle run on the same
rounded blocks with a looser bound stated in the test.  Integer outputs
(ranks, orthogonality flags) are compared exactly.
"""
import sys

import numpy as np
import pytest

import oracle as O
from oracle import philox as P
from helpers import dev_to_oracle, device_blocks_provider, rel_err_dense, to_api, to_oracle

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


def orth_ok(y, atol=1e-12):
    for k in range(y.N):
        c = y.cores[k]
        if y.ot[k] == 1:
            np.testing.assert_allclose(sum(c[i].T @ c[i] for i in range(c.shape[0])), np.eye(c.shape[2]), atol=atol)
        elif y.ot[k] == -1:
            np.testing.assert_allclose(sum(c[i] @ c[i].T for i in range(c.shape[0])), np.eye(c.shape[1]), atol=atol)


# ------------------------------------------------------------------ data paths

def test_upload_download_roundtrip(T):
    say("case: upload/download d=5")
    x = rand_tt((2, 3, 4, 3, 2), [1, 2, 5, 4, 3, 1], 1)
    dx = T.DeviceTT.upload(to_api(x))
    for k in range(x.N):
        np.testing.assert_array_equal(dx.core(k), x.cores[k])
    assert dx.rks == x.rks and dx.dims == x.dims
    dx.free()


def test_randn_matches_philox(T):
    say("case: device randn vs numpy Philox")
    dims, rks = (4, 4, 4, 4), [1, 4, 7, 4, 1]
    dx = T.DeviceTT.randn(dims, [1, 7, 7, 7, 1], seed=99)
    assert dx.rks == rks
    ref = P.rand_tt_cores(dims, rks, 99)
    for k in range(4):
        np.testing.assert_allclose(dx.core(k), ref[k], rtol=1e-14, atol=1e-15)
    dx.free()


def test_sketch_draws_match_philox(T):
    say("case: sketch draws")
    g = T.sketch_draws(5, 20000, seed=1234)
    np.testing.assert_allclose(g, P.normal_draws(1234, 5, 20000), rtol=1e-14, atol=1e-15)


# ------------------------------------------------------------------ algebra

def test_lincomb_exact(T):
    say("case: lincomb")
    dims = (3, 2, 4, 3)
    xs = [rand_tt(dims, r, 10 + r) for r in (2, 3, 4)]
    al = [0.5, -2.0, 3.0]
    out = T.lincomb([to_api(x) for x in xs], al)
    ref = O.tt_lincomb(xs, al)
    assert out.ttv_rks == ref.rks and out.ttv_ot == ref.ot
    for a, b in zip(out.ttv_vec, ref.cores):
        np.testing.assert_array_equal(a, b)
    s = to_api(xs[0]) - to_api(xs[1])
    np.testing.assert_allclose(O.ttv_to_tensor(to_oracle(s)), O.ttv_to_tensor(O.tt_sub(xs[0], xs[1])), atol=1e-13)


def test_mpo_apply(T):
    say("case: mpo apply (Hubbard L=4)")
    H = O.hubbard_1d_mpo(4, 1.0, 4.0)
    y = rand_tt(H.dims, [1, 3, 5, 3, 1], 3)
    A = T.TToperator(H.N, H.cores, H.dims, H.rks)
    out = A * to_api(y)
    ref = O.mpo_apply(H, y)
    assert out.ttv_rks == ref.rks
    for a, b in zip(out.ttv_vec, ref.cores):
        np.testing.assert_allclose(a, b, rtol=1e-14, atol=1e-14)


# ------------------------------------------------------------------ deterministic

@pytest.mark.parametrize("root", [1, 4, 7])
def test_orthogonalize(T, root):
    say(f"case: orthogonalize root={root}")
    dims = (2, 3, 2, 4, 2, 3, 2)
    x = rand_tt(dims, [1, 2, 6, 9, 12, 6, 2, 1], 20 + root)
    y = to_oracle(T.orthogonalize(to_api(x), i=root))
    ref = O.orthogonalize(x, i=root)
    assert y.rks == ref.rks and y.ot == ref.ot
    assert rel_err_dense(x, y) < 1e-13
    orth_ok(y)
    # LAPACK sign conventions: the cores agree with the oracle's, not just the tensor
    for a, b in zip(y.cores, ref.cores):
        np.testing.assert_allclose(a, b, rtol=1e-10, atol=1e-11)


def test_norm(T):
    say("case: norm")
    x = rand_tt((3, 4, 3, 4, 3), [1, 3, 8, 8, 3, 1], 5)
    assert abs(T.norm(to_api(x)) - O.tt_norm(x)) < 1e-12 * O.tt_norm(x)


@pytest.mark.parametrize("tol,rmax,direction", [(1e-12, 2 ** 14, "left"), (1e-3, 2 ** 14, "left"), (1e-12, 4, "left"),
                                                (1e-12, 2 ** 14, "right"), (1e-3, 5, "right")])
def test_tt_rounding(T, tol, rmax, direction):
    say(f"case: tt_rounding tol={tol} rmax={rmax} {direction}")
    dims = (2, 3, 4, 3, 2, 3)
    x = O.tt_add(rand_tt(dims, 4, 30), O.tt_scale(1e-3, rand_tt(dims, 5, 31)))
    y = to_oracle(T.tt_rounding(to_api(x), tol=tol, rmax=rmax, direction=direction))
    ref = O.tt_rounding(x, tol=tol, rmax=rmax, direction=direction)
    assert y.rks == ref.rks and y.ot == ref.ot
    assert O.rel_err(ref, y) < 1e-12
    orth_ok(y)


@pytest.mark.parametrize("tol", [1e-12, 1e-4])
def test_tt_rounding_large_ranks(T, tol):
    """Bond ranks 36..160: square factors of size 36..160 (1 to 5 column blocks of
    the Jacobi SVD, odd block counts included) and QR/LQ of unfoldings up to
    960 x 160 (many Householder panels)."""
    say(f"case: tt_rounding large ranks tol={tol}")
    dims = (6,) * 6
    x = O.tt_add(rand_tt(dims, 100, 33), O.tt_scale(1e-3, rand_tt(dims, 60, 34)))
    assert max(x.rks) == 160
    y = to_oracle(T.tt_rounding(to_api(x), tol=tol))
    ref = O.tt_rounding(x, tol=tol)
    assert y.rks == ref.rks and y.ot == ref.ot
    err = O.rel_err(ref, y)
    say(f"  ranks {y.rks}, rel err vs oracle {err:.3e}")
    assert err < 1e-12
    orth_ok(y)


def test_tt_rounding_closed_form(T):
    say("case: round(y + y + 1e-10 z, 1e-8) = 2y")
    dims = (4,) * 8
    yv, z = rand_tt(dims, 6, 40), rand_tt(dims, 6, 41)
    x = O.tt_add(O.tt_add(yv, yv), O.tt_scale(1e-10, z))
    r = to_oracle(T.tt_rounding(to_api(x), tol=1e-8))
    assert r.rks == yv.rks
    assert O.rel_err(O.tt_scale(2.0, yv), r) < 1e-9


def test_tt_rounding_batch_uneven(T):
    say("case: round_svd_batch with uneven ranks")
    dims = (4,) * 6
    xs = [O.tt_add(rand_tt(dims, r, 50 + r), rand_tt(dims, r + 2, 60 + r)) for r in (2, 5, 9)]
    devs = [T.DeviceTT.upload(to_api(x)) for x in xs]
    outs = T.round_svd_batch(devs, tol=1e-10)
    for x, o in zip(xs, outs):
        y = dev_to_oracle(o)
        ref = O.tt_rounding(x, tol=1e-10)
        assert y.rks == ref.rks
        assert O.rel_err(ref, y) < 1e-12
    for d in devs + outs:
        d.free()


def test_tt_rounding_errors(T):
    say("case: rounding error paths")
    x = rand_tt((2, 2, 2), 2, 1)
    with pytest.raises(ValueError):
        T.tt_rounding(to_api(x), direction="up")


# ------------------------------------------------------------------ sketch

@pytest.mark.parametrize("orthogonal", [False, True])
def test_sketch_blocks_match_generate(T, orthogonal):
    say(f"case: sketch blocks orthogonal={orthogonal}")
    dims = (3, 2, 4, 2, 3)
    rks = [1, 7, 9, 9, 5, 1]
    bv, p = O.sketch_structure(dims, rks, 4)
    for k in range(len(dims)):
        S = T.sketch_blocks(dims, rks, k, orthogonal=orthogonal, seed=77, block_rks=4)
        g = P.normal_draws(77, k, bv[k + 1] * dims[k] * bv[k] * p[k])
        ref = O.generate_sketch_blocks(g, bv[k + 1], dims[k], bv[k], p[k], orthogonal)
        np.testing.assert_allclose(S, ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 2e-6), ("bf16", 3e-2)])
@pytest.mark.parametrize("orthogonal,block_rks", [(False, None), (True, 3)])
def test_recursive_sketch(T, prec, tol, orthogonal, block_rks):
    say(f"case: recursive sketch {prec} orthogonal={orthogonal} block_rks={block_rks}")
    dims = (2, 3, 2, 4, 2, 3)
    x = rand_tt(dims, [1, 2, 6, 8, 8, 3, 1], 70)
    rks = [1, 5, 7, 7, 6, 3, 1]
    W, s = T.tt_recursive_sketch(to_api(x), rks, orthogonal=orthogonal, seed=5, block_rks=block_rks, sketch_prec=prec)
    prov = device_blocks_provider(dims, rks, orthogonal=orthogonal, seed=5, block_rks=block_rks, sketch_prec=prec)
    Wr, sr = O.tt_recursive_sketch(x, rks, prov, orthogonal, block_rks)
    assert s == sr
    for a, b in zip(W, Wr):
        assert a.shape == b.shape
        assert np.linalg.norm(a - b) <= tol * np.linalg.norm(b)


# ------------------------------------------------------------------ randomized rounding

@pytest.mark.parametrize("prec", ["f64", "f32", "bf16"])
@pytest.mark.parametrize("orthogonal", [True, False])
def test_ttrand_rounding_exact_recovery(T, prec, orthogonal):
    say(f"case: ttrand exact recovery {prec} orthogonal={orthogonal}")
    # test_ttrand.jl:67-88 pattern
    dims = (2,) * 8
    a = rand_tt(dims, [1, 2, 4, 4, 4, 4, 4, 2, 1], 80)
    pert = O.tt_add(a, O.tt_scale(1e-8, rand_tt(dims, [1, 2, 3, 3, 3, 3, 3, 2, 1], 81)))
    rks = [1, 2, 4, 8, 16, 8, 4, 2, 1]
    y = to_oracle(T.ttrand_rounding(to_api(pert), rks, orthogonal=orthogonal, sketch_prec=prec))
    assert y.ot == [1] * 7 + [0]
    # BF16 rounding of the stored sketches W (and of the core operands of the BF16
    # contractions) blurs how the 1e-8 perturbation directions are resolved; storing
    # W in BF16 alone already gives ~2e-11 in an FP64 emulation, so BF16 gets 1e-9.
    assert O.rel_err(pert, y) < (1e-9 if prec == "bf16" else 1e-11)
    orth_ok(y, atol=1e-11)


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 1e-5), ("bf16", 5e-2)])
def test_ttrand_rounding_vs_oracle_truncating(T, prec, tol):
    say(f"case: ttrand truncating vs oracle {prec}")
    dims = (3,) * 7
    ys = [rand_tt(dims, 6, 90 + j) for j in range(3)]
    al = [1.0, 0.25, 0.0625]
    rks = [1, 3, 6, 6, 6, 6, 3, 1]
    y = to_oracle(T.ttrand_rounding(al, [to_api(t) for t in ys], rks, orthogonal=False, seed=42, sketch_prec=prec))
    prov = device_blocks_provider(dims, rks, orthogonal=False, seed=42, sketch_prec=prec)
    ref = O.ttrand_rounding_lincomb(al, ys, rks, prov, False)
    assert y.rks == ref.rks
    say(f"  rel err vs oracle: {O.rel_err(ref, y):.3e}")
    assert O.rel_err(ref, y) < tol


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 1e-5)])
def test_ttrand_rounding_large_ranks(T, prec, tol):
    """Sketch ranks up to 140 with term ranks 40..90: multi-panel Householder QR of
    Z_k (up to 512 x 140) and multi-block sketch contractions."""
    say(f"case: ttrand large ranks {prec}")
    dims = (8,) * 6
    ys = [rand_tt(dims, r, 120 + j) for j, r in enumerate((90, 40, 64))]
    al = [1.0, -0.5, 0.25]
    rks = O.r_and_d_to_rks([140] * 7, dims, rmax=140)
    assert rks == [1, 8, 64, 140, 64, 8, 1]
    y = to_oracle(T.ttrand_rounding(al, [to_api(t) for t in ys], rks, orthogonal=False, seed=43, sketch_prec=prec))
    prov = device_blocks_provider(dims, rks, orthogonal=False, seed=43, sketch_prec=prec)
    ref = O.ttrand_rounding_lincomb(al, ys, rks, prov, False)
    assert y.rks == ref.rks
    err = O.rel_err(ref, y)
    say(f"  ranks {y.rks}, rel err vs oracle {err:.3e}")
    assert err < tol


# ------------------------------------------------------------------ implicit operator products

def to_api_op(H):
    import ttround as T
    return T.TToperator(H.N, [c.copy() for c in H.cores], H.dims, list(H.rks))


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 2e-6), ("bf16", 3e-2)])
@pytest.mark.parametrize("orthogonal,block_rks", [(False, None), (True, 3)])
def test_recursive_sketch_op(T, prec, tol, orthogonal, block_rks):
    """Sketch of H x without forming it vs the oracle on the device's blocks."""
    say(f"case: operator sketch {prec} orthogonal={orthogonal} block_rks={block_rks}")
    H = O.hubbard_1d_mpo(5, 1.0, 4.0)
    dims = H.dims
    x = rand_tt(dims, [1, 3, 5, 6, 4, 1], 71)
    rks = [1, 4, 9, 9, 4, 1]
    W, s = T.tt_recursive_sketch(to_api_op(H), to_api(x), rks, orthogonal=orthogonal, seed=6, block_rks=block_rks,
                                 sketch_prec=prec)
    prov = device_blocks_provider(dims, rks, orthogonal=orthogonal, seed=6, block_rks=block_rks, sketch_prec=prec)
    Wr, sr = O.tt_recursive_sketch_op(H, x, rks, prov, orthogonal, block_rks)
    assert s == sr
    for a, b in zip(W, Wr):
        assert a.shape == b.shape
        assert np.linalg.norm(a - b) <= tol * np.linalg.norm(b)


def test_ttrand_rounding_op_exact(T):
    """Sketch rank covering the rank of H y: the result represents 1.5 H y (the
    explicit product, itself pinned against the dense Hamiltonian)."""
    say("case: operator rounding exact recovery")
    H = O.hubbard_1d_mpo(4, 1.0, 4.0)
    y = rand_tt(H.dims, [1, 3, 5, 3, 1], 60)
    rks = [1, 4, 16, 4, 1]
    r = to_oracle(T.ttrand_rounding([1.5], to_api_op(H), [to_api(y)], rks, orthogonal=False, seed=9))
    err = O.rel_err(O.tt_scale(1.5, O.mpo_apply(H, y)), r)
    say(f"  rel err vs 1.5 H y: {err:.3e}")
    assert err < 1e-12
    assert r.ot == [1, 1, 1, 0]


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 1e-5), ("bf16", 5e-2)])
def test_ttrand_rounding_op_vs_oracle(T, prec, tol):
    """Truncating alpha_1 H y_1 + alpha_2 y_2 (H y_1 has rank 30) vs the oracle."""
    say(f"case: operator rounding vs oracle {prec}")
    H = O.hubbard_1d_mpo(6, 1.0, 4.0)
    dims = H.dims
    ys = [rand_tt(dims, 5, 90), rand_tt(dims, 4, 91)]
    al = [1.0, 0.5]
    rks = [1, 4, 12, 12, 12, 4, 1]
    y = to_oracle(T.ttrand_rounding(al, to_api_op(H), [to_api(t) for t in ys], rks, orthogonal=False, seed=44,
                                    sketch_prec=prec))
    prov = device_blocks_provider(dims, rks, orthogonal=False, seed=44, sketch_prec=prec)
    ref = O.ttrand_rounding_op(al, H, ys, rks, prov, False)
    assert y.rks == ref.rks
    err = O.rel_err(ref, y)
    say(f"  rel err vs oracle: {err:.3e}")
    assert err < tol


def test_ttrand_rounding_residual(T):
    """ttrand_rounding(A, y, b, rks) = rounding of A y - b, vs the oracle (truncating)
    and vs the explicit A y - b (sketch rank covering the rank)."""
    say("case: residual rounding")
    H = O.hubbard_1d_mpo(5, 1.0, 4.0)
    dims = H.dims
    y, b = rand_tt(dims, 4, 92), rand_tt(dims, 3, 93)
    rks = [1, 4, 10, 10, 4, 1]
    r = to_oracle(T.ttrand_rounding(to_api_op(H), to_api(y), to_api(b), rks, orthogonal=True, seed=45))
    prov = device_blocks_provider(dims, rks, orthogonal=True, seed=45)
    ref = O.ttrand_rounding_residual(H, y, b, rks, prov, True)
    assert r.rks == ref.rks
    assert O.rel_err(ref, r) < 1e-12
    full = [1, 4, 16, 16, 4, 1]
    r = to_oracle(T.ttrand_rounding(to_api_op(H), to_api(y), to_api(b), full, orthogonal=False, seed=45))
    assert O.rel_err(O.tt_sub(O.mpo_apply(H, y), b), r) < 1e-12


def test_round_rand_batch_mixed_operator_items(T):
    """One launch over items with and without an operator term, uneven ranks."""
    say("case: round_rand_batch with operator items")
    H = O.hubbard_1d_mpo(5, 1.0, 4.0)
    dims = H.dims
    dop = T.DeviceTTO(to_api_op(H))
    specs = [(True, (3, 2)), (False, (5, 4)), (True, (6,))]
    items, refs, devs = [], [], []
    for i, (op, rs) in enumerate(specs):
        ys = [rand_tt(dims, r, 200 + 10 * i + j) for j, r in enumerate(rs)]
        al = [1.0, -0.25][: len(ys)]
        d = [T.DeviceTT.upload(to_api(t)) for t in ys]
        devs += d
        rks = [1, 4, 11, 11, 4, 1]
        items.append((al, dop, d, rks) if op else (al, d, rks))
        prov = device_blocks_provider(dims, rks, orthogonal=False, seed=46, sketch_prec="f32")
        refs.append(O.ttrand_rounding_op(al, H, ys, rks, prov, False) if op
                    else O.ttrand_rounding_lincomb(al, ys, rks, prov, False))
    outs = T.round_rand_batch(items, orthogonal=False, seed=46, sketch_prec="f32")
    for o, ref in zip(outs, refs):
        y = dev_to_oracle(o)
        assert y.rks == ref.rks
        assert O.rel_err(ref, y) < 1e-5
    for d in devs + outs:
        d.free()
    dop.free()


def test_round_rand_batch_uneven(T):
    say("case: round_rand_batch uneven items")
    dims = (4,) * 6
    items, refs = [], []
    devs = []
    for i, r in enumerate((3, 7, 11)):
        ys = [rand_tt(dims, r + j, 100 + 10 * i + j) for j in range(2)]
        al = [1.0, 0.5]
        d = [T.DeviceTT.upload(to_api(t)) for t in ys]
        devs += d
        rks = O.default_rank_heuristic(ys)
        items.append((al, d, rks))
        prov = device_blocks_provider(dims, rks, orthogonal=False, seed=7, sketch_prec="f64")
        refs.append(O.ttrand_rounding_lincomb(al, ys, rks, prov, False))
    outs = T.round_rand_batch(items, orthogonal=False, seed=7, sketch_prec="f64")
    for o, ref in zip(outs, refs):
        y = dev_to_oracle(o)
        assert y.rks == ref.rks
        assert O.rel_err(ref, y) < 1e-12
    for d in devs + outs:
        d.free()


# ------------------------------------------------------------------ inner products

def test_dot_batch_vs_oracle(T):
    """dot (tt_operations.jl:302-321) for pairs of different lengths, dims and ranks
    in one call; |got - ref| <= 1e-12 |x| |y| (FP64, the sum may cancel)."""
    say("case: dot batch, mixed d")
    cases = [((3, 2, 4, 2, 3), [1, 3, 5, 4, 2, 1], [1, 2, 6, 3, 3, 1]),
             ((7,), [1, 1], [1, 1]),
             ((4,) * 8, [1, 4, 9, 17, 23, 17, 9, 4, 1], [1, 3, 7, 12, 30, 12, 7, 3, 1]),
             ((2, 5), [1, 2, 1], [1, 2, 1])]
    xs, ys, refs, bounds, devs = [], [], [], [], []
    for i, (dims, ra, rb) in enumerate(cases):
        x, y = rand_tt(dims, ra, 200 + i), rand_tt(dims, rb, 300 + i)
        dx, dy = T.DeviceTT.upload(to_api(x)), T.DeviceTT.upload(to_api(y))
        xs.append(dx)
        ys.append(dy)
        devs += [dx, dy]
        refs.append(O.dot(x, y))
        bounds.append(1e-12 * O.tt_norm(x) * O.tt_norm(y))
    got = T.dot_batch(xs, ys)
    for g, r, b in zip(got, refs, bounds):
        assert abs(g - r) <= b, (g, r)
    # the same pairs in the other order give the same numbers
    got2 = T.dot_batch(ys, xs)
    for g, r, b in zip(got2, refs, bounds):
        assert abs(g - r) <= b
    for d in devs:
        d.free()


def test_dot_large_ranks_equals_norm_squared(T):
    """At ranks to 200 the oracle is slow: <x, x> against |x|^2 from ttr_norm
    and <x, y> against the oracle once."""
    say("case: dot large ranks")
    dims = (4,) * 10
    x = rand_tt(dims, [1, 4, 16, 64, 200, 200, 200, 64, 16, 4, 1], 401)
    y = rand_tt(dims, [1, 4, 16, 50, 120, 150, 120, 50, 16, 4, 1], 402)
    dx, dy = T.DeviceTT.upload(to_api(x)), T.DeviceTT.upload(to_api(y))
    g = T.dot_batch([dx, dx, dx], [dx, dy, dx])
    n2 = T.norm_batch([dx])[0] ** 2
    assert abs(g[0] - n2) <= 1e-12 * n2
    assert g[0] == g[2]
    nx, ny = O.tt_norm(x), O.tt_norm(y)
    assert abs(g[1] - O.dot(x, y)) <= 1e-12 * nx * ny
    assert abs(T.dot(to_api(x), to_api(y)) - g[1]) <= 1e-13 * nx * ny
    dx.free()
    dy.free()


@pytest.mark.parametrize("reverse", [True, False])
def test_partial_contraction_vs_oracle(T, reverse):
    say(f"case: partial_contraction reverse={reverse}")
    dims = (2, 3, 2, 4, 3)
    x, y = rand_tt(dims, [1, 2, 5, 7, 3, 1], 411), rand_tt(dims, [1, 3, 4, 6, 2, 1], 412)
    W = T.partial_contraction(to_api(x), to_api(y), reverse=reverse)
    Wr = O.partial_contraction(x, y, reverse=reverse)
    scale = O.tt_norm(x) * O.tt_norm(y)
    for a, b in zip(W, Wr):
        assert a.shape == b.shape
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-12 * max(1.0, np.abs(b).max()))
    end = W[0] if reverse else W[-1]
    assert abs(end[0, 0] - O.dot(x, y)) <= 1e-12 * scale


def wide_operator(dims, R, seed):
    rng = np.random.default_rng(seed)
    Hr = [1] + [R] * (len(dims) - 1) + [1]
    return O.TTO([rng.standard_normal((n, n, Hr[k], Hr[k + 1])) / np.sqrt(n * Hr[k])
                  for k, n in enumerate(dims)], dims, Hr)


def test_dot_operator_vs_oracle(T):
    """<psi|H|phi> (tt_operations.jl:385-421): Hubbard (n R = 24, LDS-staged
    local action) and a random operator with R = 20 (n R = 80, GEMM against the
    permuted core) in one batch, against the oracle and against dot(psi, H phi)."""
    say("case: dot_operator batch (Hubbard + wide operator)")
    Hh = O.hubbard_1d_mpo(6, 1.0, 4.0)
    Hw = wide_operator((4,) * 6, 20, 421)
    trip, refs, bounds, devs = [], [], [], []
    dh, dw = T.DeviceTTO(to_api_op(Hh)), T.DeviceTTO(to_api_op(Hw))
    for i, (H, dH) in enumerate(((Hh, dh), (Hw, dw), (Hh, dh))):
        psi = rand_tt(H.dims, [1, 4, 9, 12, 9, 4, 1], 430 + i)
        phi = rand_tt(H.dims, [1, 3, 7, 10, 8, 4, 1], 440 + i)
        if i == 2:
            psi = O.orthogonalize(psi, 3)  # centre in the middle: the oracle pivots there
        dp, df = T.DeviceTT.upload(to_api(psi)), T.DeviceTT.upload(to_api(phi))
        devs += [dp, df]
        trip.append((dp, dH, df))
        refs.append(O.dot_operator(psi, H, phi))
        bounds.append(1e-12 * O.tt_norm(psi) * O.tt_norm(O.mpo_apply(H, phi)))
    got = T.dot_operator_batch([t[0] for t in trip], [t[1] for t in trip], [t[2] for t in trip])
    for g, r, b in zip(got, refs, bounds):
        say(f"  got {g:.15e} ref {r:.15e}")
        assert abs(g - r) <= b
    # implicit form equals the explicit product H phi
    hp = T.mpo_apply_device(dw, trip[1][2])
    assert abs(T.dot_batch([trip[1][0]], [hp])[0] - got[1]) <= bounds[1]
    hp.free()
    for d in devs + [dh, dw]:
        d.free()


def test_dot_operator_edge_cases(T):
    """d = 1, a rank-one operator (identity: <psi|I|phi> = dot) and the
    Hubbard energy of a normalised state against the dense Hamiltonian."""
    say("case: dot_operator edge cases")
    H1 = O.TTO([np.random.default_rng(5).standard_normal((5, 5, 1, 1))], (5,), [1, 1])
    a, b = rand_tt((5,), [1, 1], 451), rand_tt((5,), [1, 1], 452)
    ref = float(a.cores[0][:, 0, 0] @ H1.cores[0][:, :, 0, 0] @ b.cores[0][:, 0, 0])
    assert abs(T.dot_operator(to_api(a), to_api_op(H1), to_api(b)) - ref) <= 1e-13 * abs(ref) + 1e-15
    dims = (3, 4, 2, 3)
    Id = O.TTO([np.eye(n).reshape(n, n, 1, 1) for n in dims], dims, [1] * 5)
    x, y = rand_tt(dims, [1, 3, 5, 2, 1], 453), rand_tt(dims, [1, 2, 4, 3, 1], 454)
    assert abs(T.dot_operator(to_api(x), to_api_op(Id), to_api(y)) - O.dot(x, y)) <= 1e-12 * O.tt_norm(x) * O.tt_norm(y)
    Hh = O.hubbard_1d_mpo(4, 1.0, 4.0)
    v = rand_tt(Hh.dims, [1, 4, 8, 4, 1], 455)
    vd = O.ttv_to_tensor(v).reshape(-1, order="F")
    e_dense = float(vd @ O.hubbard_1d_dense(4, 1.0, 4.0) @ vd)
    e = T.dot_operator(to_api(v), to_api_op(Hh), to_api(v))
    assert abs(e - e_dense) <= 1e-12 * abs(e_dense)


def test_inner_product_errors(T):
    say("case: inner product argument errors")
    x, y = rand_tt((2, 3), [1, 2, 1], 461), rand_tt((3, 2), [1, 2, 1], 462)
    dx, dy = T.DeviceTT.upload(to_api(x)), T.DeviceTT.upload(to_api(y))
    with pytest.raises(T.DimensionMismatch):
        T.dot_batch([dx], [dy])
    with pytest.raises(T.DimensionMismatch):
        T.dot(to_api(x), to_api(y))
    dx.free()
    dy.free()


# ------------------------------------------------------------------ boundary sketches (Krylov callers)

@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 2e-6), ("bf16", 3e-2)])
@pytest.mark.parametrize("orthogonal,block_rks", [(False, None), (True, 3)])
def test_tt_sketch_batch_vs_oracle(T, prec, tol, orthogonal, block_rks):
    """tt_sketch of a batch mixing plain TTs and implicit H x products (one
    library call) vs the oracle on the device's blocks (same p at every core)."""
    say(f"case: tt_sketch batch {prec} orthogonal={orthogonal} block_rks={block_rks}")
    H = O.hubbard_1d_mpo(5, 1.0, 4.0)
    dims, N = H.dims, 5
    xs = [rand_tt(dims, [1, 3, 5, 6, 4, 1], 130 + i) for i in range(4)]
    ops = [None, H, None, H]
    s = 20
    dx = [T.DeviceTT.upload(to_api(x)) for x in xs]
    dH = T.DeviceTTO(to_api_op(H))
    S, sa = T.tt_sketch_batch(dx, s, [None if o is None else dH for o in ops], orthogonal=orthogonal, seed=8,
                              block_rks=block_rks, sketch_prec=prec)
    bv, _ = O.sketch_structure(dims, [1] * (N + 1), block_rks)
    p = -(-s // bv[0])
    prov = device_blocks_provider(dims, [1] * (N + 1), orthogonal=orthogonal, seed=8, block_rks=block_rks,
                                  sketch_prec=prec, p=[p] * N + [1])
    assert S.shape == (sa, 4) and sa == bv[0] * p
    for i, x in enumerate(xs):
        ref, sr = O.tt_sketch(x, s, prov, orthogonal, block_rks, H=ops[i])
        assert sr == sa
        err = np.linalg.norm(S[:, i] - ref) / np.linalg.norm(ref)
        say(f"  item {i} ({'H x' if ops[i] is not None else 'x'}): rel err {err:.3e}")
        assert err <= tol
    for d in dx + [dH]:
        d.free()


@pytest.mark.parametrize("op", [False, True])
@pytest.mark.parametrize("s", [8, 30])
def test_tt_combined_sketch_vs_oracle(T, op, s):
    """tt_combined_sketch(H?, x, rmax, s): s = 8 fits the recursive sketch's
    boundary rank, s = 30 needs the extra tt_sketch drawn from seed + 1."""
    say(f"case: tt_combined_sketch op={op} s={s}")
    H = O.hubbard_1d_mpo(5, 1.0, 4.0)
    dims, N = H.dims, 5
    x = rand_tt(dims, [1, 3, 5, 6, 4, 1], 140)
    rmax = 6
    args = ([to_api_op(H)] if op else []) + [to_api(x), rmax, s]
    sk, W, srk = T.tt_combined_sketch(*args, orthogonal=False, seed=9)
    rks0 = [rmax] * (N + 1)
    rks0[N] = 1
    rks = list(rks0)
    rks[0] = max(rks[0], min(s, rks[1]))
    prov = device_blocks_provider(dims, rks, orthogonal=False, seed=9)
    bv, pr = O.sketch_structure(dims, rks, None)
    bvs, _ = O.sketch_structure(dims, [1] * (N + 1), None)
    pe = max(1, -(-(s - bv[0] * pr[0]) // bvs[0]))
    prov_e = device_blocks_provider(dims, [1] * (N + 1), orthogonal=False, seed=10, p=[pe] * N + [1])
    ref, Wr, sr = O.tt_combined_sketch(x, rks0, s, prov, prov_e, False, None, H=H if op else None)
    assert srk == sr and len(sk) == len(ref) and len(sk) >= s
    assert np.linalg.norm(sk - ref) <= 1e-12 * np.linalg.norm(ref)
    for a, b in zip(W, Wr):
        assert a.shape == b.shape
        assert np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(b)


def test_tt_sketch_errors(T):
    say("case: tt_sketch argument errors")
    import ctypes as C
    from ttround import _lib as L
    x, y = rand_tt((2, 3, 2), [1, 2, 2, 1], 471), rand_tt((3, 2, 2), [1, 2, 2, 1], 472)
    dx, dy = T.DeviceTT.upload(to_api(x)), T.DeviceTT.upload(to_api(y))
    with pytest.raises(T.DimensionMismatch):
        T.tt_sketch_batch([dx, dy], 4)
    h = dx.handle
    o, _ = T.api._opts(False)
    ids = L.as_i64([dx.id])
    sa = np.zeros(1, np.int64)
    out = np.zeros(1)
    st = h.lib.ttr_tt_sketch(h.h, 1, L.ptr_i64(ids), C.cast(None, L.i64p), 4, C.byref(o), L.ptr_f64(out), 1,
                             L.ptr_i64(sa))
    assert st == L.TTR_INVALID_ARG and sa[0] >= 4
    dx.free()
    dy.free()


# ------------------------------------------------------------------ STTA (two-sided sketches)

@pytest.mark.parametrize("orthogonal,block_rks", [(False, None), (True, None), (True, 3)])
def test_forward_sketch_vs_oracle(T, orthogonal, block_rks):
    say(f"case: forward sketch orthogonal={orthogonal} block_rks={block_rks}")
    dims = (2, 3, 2, 4, 2, 3)
    x = rand_tt(dims, [1, 2, 6, 8, 8, 3, 1], 150)
    rks = [1, 5, 7, 7, 6, 3, 1]
    W, s = T.tt_recursive_sketch(to_api(x), rks, reverse=False, orthogonal=orthogonal, seed=21, block_rks=block_rks,
                                 oversampling=1.5)
    prov = device_blocks_provider(dims, rks, orthogonal=orthogonal, seed=21, block_rks=block_rks, oversampling=1.5,
                                  forward=True)
    Wr, sr = O.tt_recursive_sketch_forward(x, rks, prov, orthogonal, block_rks, oversampling=1.5)
    assert s == sr
    for a, b in zip(W, Wr):
        assert a.shape == b.shape
        assert np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(b)


def _stta_provs(dims, rks, orthogonal, block_rks, sl, sr_):
    pl = device_blocks_provider(dims, rks, orthogonal=orthogonal, seed=sl, block_rks=block_rks, oversampling=1.5,
                                forward=True)
    pr = device_blocks_provider(dims, rks, orthogonal=orthogonal, seed=sr_, block_rks=block_rks)
    return pl, pr


@pytest.mark.parametrize("orthogonal", [False, True])
def test_stta_sketch_vs_oracle(T, orthogonal):
    say(f"case: stta_sketch orthogonal={orthogonal}")
    dims = (3, 2, 4, 2, 3)
    x = rand_tt(dims, [1, 3, 5, 6, 3, 1], 151)
    rks = [1, 3, 4, 5, 3, 1]
    Om, Ps = T.stta_sketch(to_api(x), rks, seed_left=31, seed_right=32, orthogonal=orthogonal)
    pl, pr = _stta_provs(dims, rks, orthogonal, None, 31, 32)
    Omr, Psr, _, _ = O.stta_sketch(x, rks, pl, pr, orthogonal)
    for a, b in zip(Om, Omr):
        assert a.shape == b.shape and np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(b)
    for a, b in zip(Ps, Psr):
        assert a.shape == b.shape and np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(b)


@pytest.mark.parametrize("orthogonal", [False, True])
def test_stta_exact_recovery(T, orthogonal):
    # test_ttrand.jl:136-160 pattern.  With sketch ranks above the true ranks a low-rank
    # input is recovered to rounding; with a 1e-8 perturbation of higher rank the
    # pseudo-inverses (1e-12 cutoff) resolve those directions only to ~1e-7 (the FP64
    # oracle gives 4e-8..1e-6 on this case), so that case gets the reference's quality
    # bound instead (isapprox with atol 1e-6 on a unit-size tensor)
    say(f"case: stta exact recovery orthogonal={orthogonal}")
    dims = (2,) * 8
    a = rand_tt(dims, [1, 2, 4, 4, 4, 4, 4, 2, 1], 152)
    rks = [1, 2, 4, 8, 16, 8, 4, 2, 1]
    y = to_oracle(T.stta(to_api(a), rks, seed_left=7777, seed_right=8888, orthogonal=orthogonal))
    err = O.rel_err(a, y)
    say(f"  ranks {y.rks}, rel err vs input {err:.3e}")
    assert err < 1e-10 and y.ot == [0] * 8
    pl, pr = _stta_provs(dims, rks, orthogonal, None, 7777, 8888)
    ref = O.stta([1.0], [a], rks, pl, pr, orthogonal)
    assert y.rks == ref.rks and O.rel_err(ref, y) < 1e-10
    pert = O.tt_add(a, O.tt_scale(1e-8, rand_tt(dims, [1, 2, 3, 3, 3, 3, 3, 2, 1], 153)))
    yp = to_oracle(T.stta(to_api(pert), rks, seed_left=7777, seed_right=8888, orthogonal=orthogonal))
    say(f"  perturbed input: rel err {O.rel_err(pert, yp):.3e}")
    assert O.rel_err(pert, yp) < 1e-5


def test_stta_truncating_vs_oracle(T):
    say("case: stta truncating vs oracle")
    dims = (3,) * 6
    ys = [rand_tt(dims, 5, 154 + j) for j in range(3)]
    al = [1.0, 0.3, -0.1]
    rks = [1, 3, 4, 4, 4, 3, 1]
    y = to_oracle(T.stta(al, [to_api(t) for t in ys], rks, seed_left=41, seed_right=42, orthogonal=False))
    pl, pr = _stta_provs(dims, rks, False, None, 41, 42)
    ref = O.stta(al, ys, rks, pl, pr, False)
    assert y.rks == ref.rks
    err = O.rel_err(ref, y)
    say(f"  ranks {y.rks}, rel err vs oracle {err:.3e}, vs input {O.rel_err(O.tt_lincomb(ys, al), y):.3e}")
    assert err < 1e-10


def test_stta_batch_uneven(T):
    say("case: stta batch of uneven items")
    dims = (2, 3, 2, 3, 2)
    items, refs = [], []
    for i, (r, m) in enumerate(((2, 1), (4, 2), (3, 3))):
        ys = [rand_tt(dims, r, 160 + 10 * i + j) for j in range(m)]
        al = [0.5 ** j for j in range(m)]
        rks = O.default_rank_heuristic(ys)
        items.append((al, [T.DeviceTT.upload(to_api(t)) for t in ys], rks))
        pl, pr = _stta_provs(dims, rks, True, None, 51, 52)
        refs.append(O.stta(al, ys, rks, pl, pr, True))
    outs = T.stta_batch(items, seed_left=51, seed_right=52, orthogonal=True)
    for o, ref in zip(outs, refs):
        got = dev_to_oracle(o)
        assert got.rks == ref.rks
        assert O.rel_err(ref, got) < 1e-10
        o.free()
    for _, ds, _ in items:
        for d_ in ds:
            d_.free()


def test_stta_errors(T):
    say("case: stta argument errors")
    x = rand_tt((2, 3, 2), [1, 2, 2, 1], 170)
    dx = T.DeviceTT.upload(to_api(x))
    import ctypes as C
    from ttround import _lib as L
    h = dx.handle
    arr = (L.RandItem * 1)()
    ids, al, rk = L.as_i64([dx.id]), np.ones(1), L.as_i64([1, 2, 2, 1])
    arr[0].nterms, arr[0].alpha, arr[0].term_ids, arr[0].op_id, arr[0].rks = 1, L.ptr_f64(al), L.ptr_i64(ids), -1, \
        L.ptr_i64(rk)
    out = np.zeros(1, np.int64)
    o, _ = T.api._opts(False, "f32")
    assert h.lib.ttr_stta(h.h, 1, arr, C.byref(o), 5, L.ptr_i64(out)) == L.TTR_INVALID_ARG
    o, _ = T.api._opts(False)
    arr[0].op_id = 12345
    assert h.lib.ttr_stta(h.h, 1, arr, C.byref(o), 5, L.ptr_i64(out)) != L.TTR_OK
    # the randomized rounding takes the right-to-left sketch only
    arr[0].op_id = -1
    o, _ = T.api._opts(False, forward=True)
    assert h.lib.ttr_round_rand(h.h, 1, arr, C.byref(o), L.ptr_i64(out)) == L.TTR_INVALID_ARG
    dx.free()
