"""The grouped GEMM's tile configurations give bitwise identical results.

The library's GEMM runs one of eight tile configurations (64 or 128 rows and
columns, k-steps of 8 or 16), splits K into chunks of 16-element k-tiles when a
batch has few output tiles, and picks the configuration per call site: the first
call of a handle at a site times every admissible configuration (the output left
in C is the last one's), later calls run the fastest, and a handle created with
TTR_GEMM_TUNE=0 always runs 64 x 64 tiles with k-step 16.  Every output element
is accumulated in K order in all of them, so the three runs below must agree bit
for bit (FP64 and FP32 GEMMs, QR panels and trailing updates, split-K included).
"""
import os
import sys

import numpy as np
import pytest

import oracle as O
from helpers import to_api, to_oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _workload(T):
    dims = (8,) * 6
    ys = [O.rand_tt(dims, r, np.random.default_rng(120 + j)) for j, r in enumerate((90, 40, 64))]
    al = [1.0, -0.5, 0.25]
    rks = O.r_and_d_to_rks([140] * 7, dims, rmax=140)
    x = O.tt_add(O.rand_tt((6,) * 6, 100, np.random.default_rng(33)),
                 O.tt_scale(1e-3, O.rand_tt((6,) * 6, 60, np.random.default_rng(34))))
    out = []
    for prec in ("f64", "f32"):
        y = T.ttrand_rounding(al, [to_api(t) for t in ys], rks, orthogonal=False, seed=43, sketch_prec=prec)
        out.append(to_oracle(y))
    out.append(to_oracle(T.tt_rounding(to_api(x), tol=1e-10)))
    return out


def _same(a, b, what):
    for ya, yb in zip(a, b):
        assert ya.rks == yb.rks, what
        for k in range(ya.N):
            np.testing.assert_array_equal(ya.cores[k], yb.cores[k], err_msg=f"{what}: core {k}")


def test_gemm_configs_bitwise(monkeypatch):
    import ttround as T
    from ttround import _lib as L

    say("case: GEMM tile configurations, tuned handle (first and second call) vs TTR_GEMM_TUNE=0")
    tuned = L.Handle(0)
    monkeypatch.setattr(L, "_default", tuned)
    first = _workload(T)
    second = _workload(T)
    monkeypatch.setenv("TTR_GEMM_TUNE", "0")
    fixed = L.Handle(0)
    monkeypatch.setattr(L, "_default", fixed)
    plain = _workload(T)
    _same(first, second, "tuned handle, first vs second call")
    _same(first, plain, "tuned vs 64x64/k16 tiles")
    tuned.close()
    fixed.close()


@pytest.mark.parametrize("qr_path", ["0", "2"])
@pytest.mark.parametrize("cfg", range(1, 8))
def test_each_gemm_config_bitwise(monkeypatch, cfg, qr_path):
    """TTR_GEMM_CFG forces one tile configuration for every GEMM of a handle
    (split-K batches included): each must give the bits of configuration 0.
    TTR_QR_PATH=2 routes the Householder factorisations through the streaming
    panels, whose in-block updates read V and V^T through the unit-trapezoidal
    operand modes; the default path's paired blocks read an explicit V."""
    import ttround as T
    from ttround import _lib as L

    say(f"case: GEMM tile configuration {cfg} forced vs configuration 0, QR path {qr_path}")
    runs = {}
    monkeypatch.setenv("TTR_QR_PATH", qr_path)
    for c in (0, cfg):
        monkeypatch.setenv("TTR_GEMM_CFG", str(c))
        h = L.Handle(0)
        monkeypatch.setattr(L, "_default", h)
        runs[c] = _workload(T)
        h.close()
    monkeypatch.delenv("TTR_GEMM_CFG")
    monkeypatch.delenv("TTR_QR_PATH")
    _same(runs[0], runs[cfg], f"configuration {cfg} vs 0")
