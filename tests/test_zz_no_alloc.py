"""ttr_dot_noalloc (inner products without any allocation) against ttr_dot.

Two calls in a row on a fresh handle whose only earlier calls created the
inputs (so no workspace was ever allocated for an inner product on it), the
second with a smaller input, each compared with ttr_dot on the same pairs and with the FP64
oracle's dot (src/tt_operations.jl:302-321).  FP64: 1e-13 relative to |x| |y|
(both entry points run the same GEMMs; only the split of K may differ).  A call
beyond the scratch limit fails with TTR_INVALID_ARG and leaves the handle usable.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import oracle as O
from helpers import dev_to_oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _call(h, fn, xs, ys):
    from ttround import _lib as L
    a, b = L.as_i64([x.id for x in xs]), L.as_i64([y.id for y in ys])
    out = np.zeros(len(xs))
    st = getattr(h.lib, fn)(h.h, len(xs), L.ptr_i64(a), L.ptr_i64(b), L.ptr_f64(out))
    return st, out


def test_dot_noalloc_matches_dot():
    import ttround as T
    from ttround import _lib as L
    h = L.Handle(0)
    try:
        cases = [(32, 20, 64, 700), (5, 8, 16, 900)]  # (pairs, d, rank, seed base): the second is smaller
        for npair, d, r, sb in cases:
            say(f"case: dot_noalloc {npair} pairs d={d} n=4 r={r}")
            dims = (4,) * d
            xs = [T.DeviceTT.randn(dims, r, seed=sb + i, handle=h) for i in range(npair)]
            w = T.DeviceTT.randn(dims, r, seed=sb + 99, handle=h)
            ys = [w] * npair
            st, got = _call(h, "ttr_dot_noalloc", xs, ys)
            assert st == L.TTR_OK, h.lib.ttr_last_error(h.h)
            st, ref = _call(h, "ttr_dot", xs, ys)
            assert st == L.TTR_OK
            nx = T.norm_batch(xs)
            nw = T.norm_batch([w])[0]
            scale = nx * nw
            err = np.max(np.abs(got - ref) / scale)
            wo = dev_to_oracle(w)
            ora = np.array([O.dot(dev_to_oracle(x), wo) for x in xs[:4]])
            err_o = np.max(np.abs(got[:4] - ora) / scale[:4])
            say(f"  vs ttr_dot {err:.2e} (bitwise equal: {bool(np.array_equal(got, ref))}), vs oracle {err_o:.2e}")
            assert err < 1e-13 and err_o < 1e-12
            for x in xs + [w]:
                x.free()
        say("case: dot_noalloc beyond the scratch limit")
        dims = (4,) * 10
        # bond 5 has rank 512: 2 * 512^2 + 256 * 4 * 512 doubles (8.4 MB) per pair, 5 pairs > 32 MiB
        big = [T.DeviceTT.randn(dims, 512, seed=950 + i, handle=h) for i in range(5)]
        st, _ = _call(h, "ttr_dot_noalloc", big, big)
        msg = h.lib.ttr_last_error(h.h).decode()
        say(f"  status {st}: {msg}")
        assert st == L.TTR_INVALID_ARG and "MiB" in msg
        st, _ = _call(h, "ttr_dot", big[:1], big[:1])
        assert st == L.TTR_OK
        for x in big:
            x.free()
    finally:
        h.close()
