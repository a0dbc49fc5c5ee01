"""Caller-supplied sketch blocks (ttr_rand_opts.blocks / blocks_right).

A Julia caller seeds default_rng and calls the reference's generate_sketch_blocks
core by core (src/sketches.jl:44-84, 177-178, 232-237); the binding passes those
arrays and the device must use them instead of its own draws.  Here the blocks
are built by the oracle's generate_sketch_blocks from numpy draws (a stand-in for
Julia's RNG stream), passed through the C-ABI, and every result is compared with
the oracle run on the very same blocks: FP64 to 1e-12; FP32 sketches against the
oracle fed the blocks rounded to FP32, to 1e-5.  A run on the library's own draws
gives a different result, so the blocks were really used.
"""
import sys

import numpy as np
import pytest

import oracle as O
from helpers import to_api, to_oracle

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


def numpy_blocks(dims, bv, p, orthogonal, seed, forward=False):
    """Per-core blocks in generate_sketch_blocks' layout: reverse (bv_{k+1}, n_k,
    bv_k, p_k), forward (bv_k, n_k, bv_{k+1}, p_{k+1}); drawn core d..1 (reverse)
    or 1..d (forward) from one stream, as the reference's loops do."""
    g = np.random.default_rng(seed)
    d = len(dims)
    out = [None] * d
    order = range(d) if forward else range(d - 1, -1, -1)
    for k in order:
        left, right, pk = (bv[k], bv[k + 1], p[k + 1]) if forward else (bv[k + 1], bv[k], p[k])
        draws = g.standard_normal(left * dims[k] * right * pk)
        out[k] = O.generate_sketch_blocks(draws, left, dims[k], right, pk, orthogonal)
    return out


def f32(blocks):
    return [b.astype(np.float32).astype(np.float64) for b in blocks]


@pytest.mark.parametrize("orthogonal", [False, True])
def test_round_rand_with_caller_blocks(T, orthogonal):
    say(f"case: ttrand_rounding with caller blocks, orthogonal={orthogonal}")
    dims = (3, 4, 2, 4, 3, 2)
    ys = [rand_tt(dims, r, 300 + r) for r in (4, 6, 5)]
    al = [1.0, -0.4, 0.25]
    rks = [1, 3, 5, 6, 5, 3, 1]  # below the true ranks: the sketch decides the result
    bv, p = O.sketch_structure(dims, rks, None)
    blocks = numpy_blocks(dims, bv, p, orthogonal, seed=11)
    y = to_oracle(T.ttrand_rounding(al, [to_api(t) for t in ys], rks, orthogonal=orthogonal, blocks=blocks))
    ref = O.ttrand_rounding_lincomb(al, ys, rks, lambda k: blocks[k], orthogonal)
    err = O.rel_err(ref, y)
    own = to_oracle(T.ttrand_rounding(al, [to_api(t) for t in ys], rks, orthogonal=orthogonal, seed=11))
    diff = O.rel_err(own, y)
    say(f"  rel err vs oracle on the same blocks {err:.3e}; library draws differ by {diff:.3e}")
    assert y.rks == ref.rks and err < 1e-12
    assert diff > 1e-6
    # FP32 sketch contractions on the same blocks
    y32 = to_oracle(T.ttrand_rounding(al, [to_api(t) for t in ys], rks, orthogonal=orthogonal, blocks=blocks,
                                      sketch_prec="f32"))
    b32 = f32(blocks)
    ref32 = O.ttrand_rounding_lincomb(al, ys, rks, lambda k: b32[k], orthogonal)
    e32 = O.rel_err(ref32, y32)
    say(f"  FP32 sketch: rel err vs oracle on the FP32-rounded blocks {e32:.3e}")
    assert y32.rks == ref32.rks and e32 < 1e-5


def test_operator_rounding_with_caller_blocks(T):
    say("case: ttrand_rounding(alpha, H, y) with caller blocks")
    H = O.hubbard_1d_mpo(4, 1.0, 4.0)
    dims = H.dims
    ys = [rand_tt(dims, [1, 3, 5, 3, 1], 320), rand_tt(dims, [1, 2, 4, 2, 1], 321)]
    al = [1.0, -0.7]
    rks = [1, 3, 6, 3, 1]
    bv, p = O.sketch_structure(dims, rks, None)
    blocks = numpy_blocks(dims, bv, p, False, seed=12)
    Hh = T.TToperator(H.N, [c.copy() for c in H.cores], H.dims, list(H.rks))
    y = to_oracle(T.ttrand_rounding(al, Hh, [to_api(t) for t in ys], rks, orthogonal=False, blocks=blocks))
    ref = O.ttrand_rounding_op(al, H, ys, rks, lambda k: blocks[k], False)
    err = O.rel_err(ref, y)
    say(f"  rel err vs oracle {err:.3e}")
    assert y.rks == ref.rks and err < 1e-12


def test_recursive_sketches_with_caller_blocks(T):
    say("case: tt_recursive_sketch (both directions) and tt_sketch with caller blocks")
    dims = (2, 3, 4, 3, 2)
    x = rand_tt(dims, [1, 2, 5, 6, 2, 1], 330)
    rks = [1, 3, 4, 4, 3, 1]
    bv, p = O.sketch_structure(dims, rks, 2)
    blocks = numpy_blocks(dims, bv, p, True, seed=13)
    W, srk = T.tt_recursive_sketch(to_api(x), rks, orthogonal=True, block_rks=2, blocks=blocks)
    Wr, sr = O.tt_recursive_sketch(x, rks, lambda k: blocks[k], True, 2)
    assert list(srk) == list(sr)
    for a, b in zip(W, Wr):
        assert np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(b)
    rf = [1, 3, 4, 4, 3, 1]
    bvf, pf = O.sketch_structure_forward(dims, rf, 2)
    fb = numpy_blocks(dims, bvf, pf, False, seed=14, forward=True)
    Wf, sf = T.tt_recursive_sketch(to_api(x), rf, reverse=False, orthogonal=False, block_rks=2, blocks=fb)
    Wfr, sfr = O.tt_recursive_sketch_forward(x, rf, lambda k: fb[k], False, 2)
    assert list(sf) == list(sfr)
    for a, b in zip(Wf, Wfr):
        assert np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(b)
    # tt_sketch: the same block count p = cld(s, bv_1) at every core
    s = 7
    bvs, _ = O.sketch_structure(dims, [1] * 6, None)
    ps = -(-s // bvs[0])
    sb = numpy_blocks(dims, bvs, [ps] * 5 + [1], False, seed=15)
    dev = T.DeviceTT.upload(to_api(x))
    got, sa = T.tt_sketch_batch([dev], s, orthogonal=False, blocks=sb)
    dev.free()
    ref, sar = O.tt_sketch(x, s, lambda k: sb[k], False)
    assert sa == sar
    assert np.linalg.norm(got[:, 0] - ref) <= 1e-12 * np.linalg.norm(ref)
    say("  reverse, forward and boundary sketches equal the oracle's on the caller's blocks")


def test_stta_with_caller_blocks(T):
    say("case: stta with caller blocks for both sketches")
    dims = (3,) * 6
    ys = [rand_tt(dims, 5, 340 + j) for j in range(2)]
    al = [1.0, 0.3]
    rks = [1, 3, 4, 4, 4, 3, 1]
    bvl, pl = O.sketch_structure_forward(dims, rks, None, oversampling=1.5)
    bvr, pr = O.sketch_structure(dims, rks, None)
    bl = numpy_blocks(dims, bvl, pl, False, seed=16, forward=True)
    br = numpy_blocks(dims, bvr, pr, False, seed=17)
    y = to_oracle(T.stta(al, [to_api(t) for t in ys], rks, orthogonal=False, blocks_left=bl, blocks_right=br))
    ref = O.stta(al, ys, rks, lambda k: bl[k], lambda k: br[k], False)
    err = O.rel_err(ref, y)
    say(f"  rel err vs oracle {err:.3e}")
    assert y.rks == ref.rks and err < 1e-10


def test_caller_blocks_need_equal_block_counts(T):
    from ttround import _lib as L
    say("case: a batch with caller blocks and different block counts is rejected")
    dims = (2, 2, 2, 2)
    a = T.DeviceTT.upload(to_api(rand_tt(dims, 2, 350)))
    r1, r2 = [1, 2, 4, 2, 1], [1, 2, 8, 2, 1]
    bv, p = O.sketch_structure(dims, r2, 1)
    blocks = numpy_blocks(dims, bv, p, False, seed=18)
    with pytest.raises(ValueError, match="same block counts"):
        T.round_rand_batch([([1.0], [a], r1), ([1.0], [a], r2)], orthogonal=False, block_rks=1, blocks=blocks)
    a.free()
    assert L.default_handle() is not None


@pytest.mark.parametrize("prec", ["f32", "bf16"])
def test_sketch_accum_f64_is_the_oracle_on_rounded_blocks(T, prec):
    """sketch_accum="f64": only the blocks are rounded to the sketch precision, the
    contractions run in FP64, which is exactly what the oracle computes on the
    rounded blocks: FP64 agreement (1e-12) where FP32/BF16 accumulation does not."""
    say(f"case: {prec} blocks with FP64 accumulation vs the oracle on the same rounded blocks")
    dims = (4,) * 10
    ys = [rand_tt(dims, r, 360 + r) for r in (12, 20, 9)]
    al = [1.0, 0.5, 0.25]
    rks = [1, 4, 16, 24, 24, 24, 24, 24, 16, 4, 1]  # below the true ranks
    got = {}
    for acc in ("f64", "sketch"):
        got[acc] = to_oracle(T.ttrand_rounding(al, [to_api(t) for t in ys], rks, orthogonal=False, seed=77,
                                               sketch_prec=prec, sketch_accum=acc))
    blocks = {k: T.sketch_blocks(dims, rks, k, orthogonal=False, seed=77, sketch_prec=prec) for k in range(10)}
    ref = O.ttrand_rounding_lincomb(al, ys, rks, blocks.__getitem__, False)
    e64, es = O.rel_err(ref, got["f64"]), O.rel_err(ref, got["sketch"])
    say(f"  rel err vs oracle: FP64 accumulation {e64:.3e}, {prec} accumulation {es:.3e}")
    assert got["f64"].rks == ref.rks and e64 < 1e-12
