"""The benchmark's five input configurations at their full sizes, through the
C-ABI, against the FP64 oracle run on the same inputs and the device's own
sketch blocks (upcast to FP64; for FP32/BF16 sketches the rounded values).

  C0  round(y + y + 1e-10 z; tol 1e-8), d=16, n=4, r=30: ranks of y, the
      oracle's result to 1e-12, and the closed form 2y to the size of the
      dropped 1e-10 z term
  C1  sum_k 2^-k y_k, 6 TTs d=32 n=4 r=100, FP64 Gaussian TT sketch of rank 144
      (block rank 144), then tt_rounding(rmax=128): both the randomized stage
      and the truncated result to 1e-12 (two FP64 LAPACK drivers, gesdd and
      gesvd, agree to about 3e-14 on this input; the smallest relative gap at a
      rank-128 cut is 0.0037)
  C2  quantics d=64 n=2, 8 TTs r=128 weighted 10^(-k/2), sketch rank 272, FP32
      and BF16 sketch contractions, cleanup rmax 256
  C4  Hubbard MPO (L=32, t=1, U=4 in the textbook convention U n_up n_dn) times
      64 TTs r=128 without forming the products, sketch rank 144, cleanup rmax
      128: items 0, 21, 42 and 63 of the batch, FP32 and BF16 sketches
(C3, the 256-item batch, is checked at full size by test_gpu_fullsize and by the
benchmark's parity leg.)  Relative errors are Frobenius norms of the represented
tensors, computed through the QR-based TT norm of the difference.
"""
import sys
import time

import numpy as np
import pytest

import oracle as O
from helpers import dev_to_oracle, to_api

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def T():
    import ttround
    ttround.default_handle()
    return ttround


def _blocks(T, dims, rks, block_rks, prec):
    cache = {}

    def get(k):
        if k not in cache:
            cache[k] = T.sketch_blocks(dims, rks, k, orthogonal=False, seed=1234, block_rks=block_rks,
                                       sketch_prec=prec)
        return cache[k]
    return get


def test_c0_fp64_svd_rounding(T):
    say("case: C0 round(y + y + 1e-10 z, tol=1e-8), d=16 n=4 r=30")
    dims = (4,) * 16
    y = dev_to_oracle(T.DeviceTT.randn(dims, 30, seed=1))
    z = dev_to_oracle(T.DeviceTT.randn(dims, 30, seed=2))
    x = O.tt_add(O.tt_add(y, y), O.tt_scale(1e-10, z))
    assert max(x.rks) == 90
    dev = T.DeviceTT.upload(to_api(x))
    out = T.round_svd_batch([dev], tol=1e-8)[0]
    r = dev_to_oracle(out)
    ref = O.tt_rounding(x, tol=1e-8)
    err = O.rel_err(ref, r)
    err2 = O.rel_err(O.tt_scale(2.0, y), r)
    say(f"  ranks {max(r.rks)} (y: {max(y.rks)}), rel err vs oracle {err:.3e}, vs 2y {err2:.3e}")
    assert r.rks == y.rks == ref.rks
    assert err < 1e-12
    # the dropped term: |1e-10 z| / |2y| (about 5e-11 for equal-rank N(0,1) TTs)
    assert err2 < 10 * 1e-10 * O.tt_norm(z) / O.tt_norm(O.tt_scale(2.0, y))
    dev.free()
    out.free()


def _c1_inputs(T):
    d, n, r = 32, 4, 100
    dims = (n,) * d
    devs = [T.DeviceTT.randn(dims, r, seed=1000 + j) for j in range(1, 7)]
    al = [2.0 ** -j for j in range(1, 7)]
    rks = [1] + [144] * (d - 1) + [1]
    return dims, devs, al, rks


def test_c1_fp64_randomized_and_cleanup(T):
    say("case: C1 6 x r=100, d=32 n=4, FP64 sketch 144, cleanup rmax=128")
    dims, devs, al, rks = _c1_inputs(T)
    raw = T.round_rand_batch([(al, devs, rks)], orthogonal=False, seed=1234, block_rks=144, sketch_prec="f64")
    clean = T.round_svd_batch(raw, rmax=128)
    ys = [dev_to_oracle(t) for t in devs]
    t0 = time.time()
    ref_raw = O.ttrand_rounding_lincomb(al, ys, rks, _blocks(T, dims, rks, 144, "f64"), False, 144)
    ref = O.tt_rounding(ref_raw, rmax=128)
    got_raw, got = dev_to_oracle(raw[0]), dev_to_oracle(clean[0])
    e_raw, e = O.rel_err(ref_raw, got_raw), O.rel_err(ref, got)
    x = O.tt_lincomb(ys, al)
    e_in = O.rel_err(x, got)
    say(f"  oracle {time.time() - t0:.1f}s; randomized stage {e_raw:.3e}, after cleanup {e:.3e}, "
        f"vs input {e_in:.3e}; ranks {max(got.rks)}")
    assert got_raw.rks == ref_raw.rks and got.rks == ref.rks
    assert e_raw < 1e-12
    assert e < 1e-12
    for t in devs + raw + clean:
        t.free()


@pytest.mark.parametrize("prec,bound", [("f32", 1e-5), ("bf16", 5e-2)])
def test_c2_quantics_mixed_precision(T, prec, bound):
    say(f"case: C2 quantics d=64 n=2, 8 x r=128, {prec} sketch 272, cleanup rmax=256")
    d = 64
    dims = (2,) * d
    devs = [T.DeviceTT.randn(dims, 128, seed=2000 + j) for j in range(1, 9)]
    al = [10.0 ** (-j / 2) for j in range(1, 9)]
    rks = [1] + [272] * (d - 1) + [1]
    raw = T.round_rand_batch([(al, devs, rks)], orthogonal=False, seed=1234, block_rks=272, sketch_prec=prec)
    clean = T.round_svd_batch(raw, rmax=256)
    ys = [dev_to_oracle(t) for t in devs]
    t0 = time.time()
    ref_raw = O.ttrand_rounding_lincomb(al, ys, rks, _blocks(T, dims, rks, 272, prec), False, 272)
    ref = O.tt_rounding(ref_raw, rmax=256)
    got_raw, got = dev_to_oracle(raw[0]), dev_to_oracle(clean[0])
    e_raw, e = O.rel_err(ref_raw, got_raw), O.rel_err(ref, got)
    say(f"  oracle {time.time() - t0:.1f}s; randomized stage {e_raw:.3e}, after cleanup {e:.3e}; "
        f"ranks {max(got.rks)}")
    assert got.rks == ref.rks
    assert e_raw < bound and e < bound
    for t in devs + raw + clean:
        t.free()


@pytest.mark.parametrize("prec,bound", [("f32", 1e-5), ("bf16", 5e-2)])
def test_c4_hubbard_batch(T, prec, bound):
    say(f"case: C4 Hubbard L=32 MPO x 64 TTs r=128, {prec} sketch 144, cleanup rmax=128")
    L = 32
    dims = (4,) * L
    Hh = T.hubbard_1d_mpo(L, 1.0, 4.0)
    H = T.DeviceTTO(Hh)
    devs = [T.DeviceTT.randn(dims, 128, seed=4001 + 100000 * b) for b in range(64)]
    rks = [1] + [144] * (L - 1) + [1]
    batch = [([1.0], H, [t], rks) for t in devs]
    raw = T.round_rand_batch(batch, orthogonal=False, seed=1234, block_rks=144, sketch_prec=prec)
    clean = T.round_svd_batch(raw, rmax=128)
    Ho = O.TTO([np.asarray(c) for c in Hh.tto_vec], Hh.tto_dims, list(Hh.tto_rks))
    blocks = _blocks(T, dims, rks, 144, prec)
    worst = 0.0
    t0 = time.time()
    for b in (0, 21, 42, 63):
        y = dev_to_oracle(devs[b])
        ref_raw = O.ttrand_rounding_op([1.0], Ho, [y], rks, blocks, False, 144)
        ref = O.tt_rounding(ref_raw, rmax=128)
        got = dev_to_oracle(clean[b])
        e_raw, e = O.rel_err(ref_raw, dev_to_oracle(raw[b])), O.rel_err(ref, got)
        say(f"  item {b}: randomized stage {e_raw:.3e}, after cleanup {e:.3e}, ranks {max(got.rks)}")
        assert got.rks == ref.rks
        worst = max(worst, e_raw, e)
    say(f"  oracle {time.time() - t0:.1f}s for 4 items, worst {worst:.3e}")
    assert worst < bound
    for t in devs + raw + clean:
        t.free()
    H.free()
