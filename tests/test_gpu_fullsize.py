"""Full-size checks of the batched randomized rounding (the C3 shape: sums of 8
N(0,1) TTs, d=40, n=4, ranks U{16..256}, coefficients 0.5^j; FP32 sketch; FP64
SVD cleanup tol=1e-6) through properties that do not need the oracle at this
size: the reference's output ranks (min(n r_{k-1}, sketch rank)), left-orthonormal cores after the randomized stage,
right-orthonormal cores after the cleanup, the randomized stage as the
orthogonal projection of the exact sum onto its frame, and the cleanup within
the TT-SVD error bound of its tolerance (errors from device inner products)."""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def T():
    import ttround
    ttround.default_handle()
    return ttround


def _gram_err(c, axis_in):
    # max |sum_i c_i^T c_i - I| over the cores' orthonormal side
    n, r, rp = c.shape
    if axis_in == "left":
        M = sum(c[i].T @ c[i] for i in range(n))
    else:
        M = sum(c[i] @ c[i].T for i in range(n))
    return float(np.abs(M - np.eye(M.shape[0])).max())


def test_c3_items_full_size(T):
    say("case: two C3-shaped items (d=40, n=4, 8 terms r~U{16..256}), FP32 sketch + cleanup tol=1e-6")
    rng = np.random.default_rng(3)
    d, n = 40, 4
    dims = (n,) * d
    items, sums = [], []
    for b in range(2):
        terms, alphas, rk = [], [], []
        for j in range(1, 9):
            r = int(rng.integers(16, 257))
            terms.append(T.DeviceTT.randn(dims, r, seed=7000 + j + 1000 * b))
            alphas.append(0.5 ** j)
            rk.append(terms[-1].rks)
        rks = [max(q[k] for q in rk) for k in range(d + 1)]
        ovs = int(math.ceil(0.5 * max(rks)))
        rks = [1] + [v + ovs for v in rks[1:d]] + [1]
        items.append((alphas, terms, rks))
        sums.append(T.lincomb_device(alphas, terms))
    outs = T.round_rand_batch(items, orthogonal=False, seed=1234, block_rks=None, sketch_prec="f32")
    clean = T.round_svd_batch(outs, tol=1e-6)
    for b, (o, c, x) in enumerate(zip(outs, clean, sums)):
        # output bond k: min(n r_{k-1}, sketch rank bv_k p_k), the sketch rank covering the target
        bv, pk = T.block_structure(dims, items[b][2])
        for k in range(1, d):
            assert bv[k] * pk[k] >= items[b][2][k]
            assert o.rks[k] == min(n * o.rks[k - 1], bv[k] * pk[k]), (k, o.rks[k])
        assert o.ot == [1] * (d - 1) + [0]
        worst = max(_gram_err(o.core(k), "left") for k in range(d - 1))
        assert worst < 1e-10, worst
        assert c.ot == [0] + [-1] * (d - 1)
        worst = max(_gram_err(c.core(k), "right") for k in range(1, d))
        assert worst < 1e-10, worst
        # the randomized stage is the orthogonal projection of x onto the span of its
        # left-orthonormal frame (last core = frame^T x): <x, o> = <o, o>
        xx, xo, oo = T.dot_batch([x, x, o], [x, o, o])
        assert abs(xo - oo) <= 1e-9 * xx, (xo, oo, xx)
        err_rand = math.sqrt(max(xx - oo, 0.0) / xx)
        # the cleanup's floor rule drops at most tol^2 of the energy per bond:
        # ||o - c|| <= sqrt(d - 1) tol ||o|| (TT-SVD bound)
        oc, cc = T.dot_batch([o, c], [c, c])
        err_clean = math.sqrt(max(oo - 2 * oc + cc, 0.0) / oo)
        say(f"  item {b}: ranks max {max(o.rks)} -> {max(c.rks)}, truncation error of the randomized "
            f"stage {err_rand:.2e}, cleanup change {err_clean:.2e}")
        assert 0.0 < err_rand < 0.5
        assert err_clean <= math.sqrt(d - 1) * 1e-6 * 1.01
    for dd in [t for it in items for t in it[1]] + outs + clean + sums:
        dd.free()
