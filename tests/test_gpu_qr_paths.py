"""The three Householder QR paths against the oracle (src/tt_rounding.jl:183-215
orthogonalization and SVD sweeps, src/tt_randtools.jl:125-151 randomized
projection, both through LAPACK's dgeqrf/dorgqr).

The library picks a QR path by size and batch: one launch per batch for batches of
at least 256 matrices of at most 1024 rows (or of single 16-column panels), else the
blocked path (one launch per 64-column block with the 16-column panels and their
in-block updates on chip, up to 3072 rows), else the blocked path with streaming
32-column panels and grouped-GEMM in-block updates.  A handle made with
TTR_QR_PATH = 1 (blocked at every size), 2 (streaming panels at every size) or 3 (one
launch wherever the rows fit) runs the same roundings through the other paths; each must equal the oracle to
1e-12 (represented tensor, relative Frobenius norm), and the profile shows which
kernel ran.  Shapes: unfoldings of 216 x 160 (three blocks, the last one two
panels) and 960 x 72 (a last block of 8 columns); of 150 columns (a last block of
22 columns: a panel of 6); of 2080 x 128 (five rows per thread of the block kernel,
the blocked path at every setting); and randomized projections whose Z_k reaches
1120 rows.
"""
import sys

import numpy as np
import pytest

import oracle as O
from helpers import device_blocks_provider, to_api, to_oracle

pytestmark = pytest.mark.gpu

PATH_KERNEL = {"0": None, "1": "qr_block", "2": "qr_panel", "3": "qr_small"}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rand_tt(dims, rks, seed):
    return O.rand_tt(dims, rks, np.random.default_rng(seed))


def handle(monkeypatch, path):
    from ttround import _lib as L
    monkeypatch.setenv("TTR_QR_PATH", path)
    h = L.Handle(0)
    monkeypatch.delenv("TTR_QR_PATH")
    monkeypatch.setattr(L, "_default", h)
    h.profile(True)
    return h


def check_kernel(h, path):
    st = h.profile_stats()
    k = PATH_KERNEL[path]
    if k is not None:
        assert st.get(k, {"launches": 0})["launches"] > 0, (path, sorted(st))
    if path == "1":
        assert "qr_small" not in st and "qr_panel" not in st
    if path == "2":
        assert "qr_small" not in st and "qr_block" not in st


@pytest.mark.parametrize("path", ["0", "1", "2", "3"])
@pytest.mark.parametrize("dims,r1,r2,tol", [((6,) * 6, 100, 60, 1e-12), ((6,) * 6, 100, 60, 1e-4),
                                            ((5,) * 6, 100, 50, 1e-12), ((8,) * 6, 160, 100, 1e-8)])
def test_tt_rounding_qr_paths(monkeypatch, path, dims, r1, r2, tol):
    import ttround as T
    x = O.tt_add(rand_tt(dims, r1, 33 + r1), O.tt_scale(1e-3, rand_tt(dims, r2, 34 + r2)))
    say(f"case: tt_rounding QR path {path}, dims {dims}, ranks {x.rks}, tol {tol}")
    ref = O.tt_rounding(x, tol=tol)
    h = handle(monkeypatch, path)
    try:
        y = to_oracle(T.tt_rounding(to_api(x), tol=tol))
        check_kernel(h, path)
    finally:
        h.close()
    err = O.rel_err(ref, y)
    say(f"  ranks {y.rks}, rel err vs oracle {err:.3e}")
    assert y.rks == ref.rks and y.ot == ref.ot
    assert err < 1e-12


@pytest.mark.parametrize("path", ["0", "1", "2", "3"])
def test_ttrand_rounding_qr_paths(monkeypatch, path):
    import ttround as T
    dims = (8,) * 6
    ys = [rand_tt(dims, r, 120 + j) for j, r in enumerate((90, 40, 64))]
    al = [1.0, -0.5, 0.25]
    rks = O.r_and_d_to_rks([140] * 7, dims, rmax=140)
    say(f"case: ttrand_rounding QR path {path}, sketch ranks {rks} (Z_k up to {max(a * 8 for a in rks)} rows)")
    prov = device_blocks_provider(dims, rks, orthogonal=False, seed=43, sketch_prec="f64")
    ref = O.ttrand_rounding_lincomb(al, ys, rks, prov, False)
    h = handle(monkeypatch, path)
    try:
        y = to_oracle(T.ttrand_rounding(al, [to_api(t) for t in ys], rks, orthogonal=False, seed=43,
                                        sketch_prec="f64"))
        check_kernel(h, path)
    finally:
        h.close()
    err = O.rel_err(ref, y)
    say(f"  ranks {y.rks}, rel err vs oracle {err:.3e}")
    assert y.rks == ref.rks
    assert err < 1e-12


@pytest.mark.parametrize("group", ["1", "4"])
def test_qr_groups_vs_oracle(monkeypatch, group):
    """TTR_QR_GROUP sets how many 64-column blocks of the blocked QR form one compact
    WY group (default 2): single blocks and groups of four, with the blocked path forced,
    round to the oracle's result at 1e-12 (unfoldings of up to 2592 rows and 432
    columns: seven blocks, a group of four and a partial group of three)."""
    import ttround as T
    dims = (6,) * 6
    x = O.tt_add(rand_tt(dims, 216, 71), O.tt_scale(1e-3, rand_tt(dims, 216, 72)))
    ys = [rand_tt(dims, r, 130 + j) for j, r in enumerate((150, 40, 64))]
    al = [1.0, -0.5, 0.25]
    rks = O.r_and_d_to_rks([216] * 7, dims, rmax=216)
    say(f"case: QR groups of {group} blocks, tt_rounding ranks {x.rks} and ttrand_rounding sketch ranks {rks}")
    ref = O.tt_rounding(x, tol=1e-10)
    prov = device_blocks_provider(dims, rks, orthogonal=False, seed=44, sketch_prec="f64")
    ref2 = O.ttrand_rounding_lincomb(al, ys, rks, prov, False)
    monkeypatch.setenv("TTR_QR_GROUP", group)
    h = handle(monkeypatch, "1")
    monkeypatch.delenv("TTR_QR_GROUP")
    try:
        y = to_oracle(T.tt_rounding(to_api(x), tol=1e-10))
        y2 = to_oracle(T.ttrand_rounding(al, [to_api(t) for t in ys], rks, orthogonal=False, seed=44,
                                         sketch_prec="f64"))
        check_kernel(h, "1")
    finally:
        h.close()
    err, err2 = O.rel_err(ref, y), O.rel_err(ref2, y2)
    say(f"  ranks {y.rks} / {y2.rks}, rel err vs oracle {err:.3e} / {err2:.3e}")
    assert y.rks == ref.rks and y2.rks == ref2.rks
    assert err < 1e-12 and err2 < 1e-12
