"""ttr_round_rand through cached HIP graphs (ttr_graph_enable).

Per call shape the first call runs eagerly, the second is captured into a graph
and launched, later calls replay it with new TT values, coefficients, addresses
and (caller-supplied) sketch blocks written into the pinned staging the graph's
copy nodes read.  Every call of a graph handle is compared bit for bit with the
same call on a handle without graphs; eager work (deterministic rounding, norms,
downloads) runs on the graph handle between the replays.  GEMM tiles are chosen
by size on both handles (TTR_GEMM_TUNE=0) so that both run the same kernels.
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


def make_handle(monkeypatch, graphs):
    from ttround import _lib as L
    monkeypatch.setenv("TTR_GEMM_TUNE", "0")
    h = L.Handle(0)
    monkeypatch.delenv("TTR_GEMM_TUNE")
    h.graphs(graphs)
    return h


def rounding(T, h, items, rks, keep=None, **kw):
    """items: [(alphas, [oracle TT...], op or None)]; uploads to h, rounds, downloads."""
    devs, ops, batch = [], [], []
    for al, terms, op in items:
        ds = [T.DeviceTT.upload(to_api(t), handle=h) for t in terms]
        devs += ds
        if op is None:
            batch.append((al, ds, rks))
        else:
            H = T.DeviceTTO(T.TToperator(op.N, [c.copy() for c in op.cores], op.dims, list(op.rks)), handle=h)
            ops.append(H)
            batch.append((al, H, ds, rks))
    outs = T.round_rand_batch(batch, **kw)
    res = [to_oracle(o.to_host()) for o in outs]
    if keep is not None:  # outputs and inputs kept alive: later calls get other addresses
        keep += outs + devs
    else:
        for o in outs + devs:
            o.free()
    for H in ops:
        H.free()
    return res


def identical(a, b):
    return a.rks == b.rks and all(np.array_equal(x, y) for x, y in zip(a.cores, b.cores))


@pytest.mark.parametrize("orthogonal", [False, True])
def test_replays_equal_eager_calls(monkeypatch, orthogonal):
    import ttround as T
    dims = (4,) * 8
    rks = [1, 4, 10, 14, 14, 14, 10, 4, 1]
    hg, he = make_handle(monkeypatch, True), make_handle(monkeypatch, False)
    keep = []
    try:
        for call in range(5):
            rng = np.random.default_rng(900 + call)
            items = [([1.0 + 0.1 * call, -0.5], [O.rand_tt(dims, 7, rng), O.rand_tt(dims, 5, rng)], None),
                     ([0.3 * (call + 1)], [O.rand_tt(dims, 9, rng)], None),
                     ([2.0, 1.0, -1.0], [O.rand_tt(dims, r, rng) for r in (3, 4, 6)], None)]
            say(f"case: graphs orthogonal={orthogonal} call {call}")
            got = rounding(T, hg, items, rks, keep=keep if call % 2 == 0 else None, orthogonal=orthogonal, seed=5)
            ref = rounding(T, he, items, rks, orthogonal=orthogonal, seed=5)
            for g, r in zip(got, ref):
                assert identical(g, r)
            if call == 0:
                blocks = {k: T.sketch_blocks(dims, rks, k, orthogonal=orthogonal, seed=5) for k in range(8)}
                for (al, terms, _), g in zip(items, got):
                    o = O.ttrand_rounding_lincomb(al, terms, rks, blocks.__getitem__, orthogonal)
                    assert o.rks == g.rks and O.rel_err(o, g) < 1e-12
            # eager work on the graph handle between the replays
            x = T.DeviceTT.upload(to_api(O.rand_tt(dims, 6, rng)), handle=hg)
            y = T.round_svd_batch([x], tol=1e-8)[0]
            assert abs(y.norm() - x.norm()) <= 1e-8 * x.norm()
            x.free()
            y.free()
        st = hg.graph_stats()
        say(f"  graph stats {st}")
        assert st == {"captures": 1, "replays": 3, "fallbacks": 0, "cached": 1}
        assert he.graph_stats()["captures"] == 0
    finally:
        for t in keep:
            t.free()
        hg.close()
        he.close()


def test_operator_blocks_and_fp32(monkeypatch):
    """An operator item, FP32 sketches and caller blocks that change every call
    (same shape): the blocks travel through the pinned staging."""
    import ttround as T
    H = O.hubbard_1d_mpo(4, 1.0, 4.0)
    dims = H.dims
    rks = [1, 3, 6, 3, 1]
    bv, p = O.sketch_structure(dims, rks, None)
    hg, he = make_handle(monkeypatch, True), make_handle(monkeypatch, False)
    try:
        for call in range(4):
            rng = np.random.default_rng(950 + call)
            g = np.random.default_rng(60 + call)
            blocks = [None] * 4
            for k in range(3, -1, -1):
                draws = g.standard_normal(bv[k + 1] * dims[k] * bv[k] * p[k])
                blocks[k] = O.generate_sketch_blocks(draws, bv[k + 1], dims[k], bv[k], p[k], False)
            items = [([1.0, -0.7], [O.rand_tt(dims, [1, 3, 5, 3, 1], rng), O.rand_tt(dims, [1, 2, 4, 2, 1], rng)], H),
                     ([0.5], [O.rand_tt(dims, [1, 4, 6, 4, 1], rng)], None)]
            say(f"case: graphs with an operator item, FP32 sketch, caller blocks, call {call}")
            got = rounding(T, hg, items, rks, orthogonal=False, blocks=blocks, sketch_prec="f32")
            ref = rounding(T, he, items, rks, orthogonal=False, blocks=blocks, sketch_prec="f32")
            for a, b in zip(got, ref):
                assert identical(a, b)
        st = hg.graph_stats()
        say(f"  graph stats {st}")
        assert st["captures"] == 1 and st["replays"] == 2 and st["fallbacks"] == 0
    finally:
        hg.close()
        he.close()


def test_shapes_are_keyed_and_graphs_can_be_dropped(monkeypatch):
    """Two shapes alternate: each gets its own graph; turning graphs off frees them."""
    import ttround as T
    hg, he = make_handle(monkeypatch, True), make_handle(monkeypatch, False)
    try:
        shapes = [((3,) * 6, [1, 3, 6, 6, 6, 3, 1]), ((5,) * 5, [1, 5, 8, 8, 5, 1])]
        for call in range(6):
            dims, rks = shapes[call % 2]
            rng = np.random.default_rng(980 + call)
            items = [([1.0, 0.25], [O.rand_tt(dims, 4, rng), O.rand_tt(dims, 6, rng)], None)]
            say(f"case: graphs over two alternating shapes, call {call}")
            got = rounding(T, hg, items, rks, orthogonal=False, seed=9)
            ref = rounding(T, he, items, rks, orthogonal=False, seed=9)
            assert identical(got[0], ref[0])
        st = hg.graph_stats()
        say(f"  graph stats {st}")
        assert st == {"captures": 2, "replays": 2, "fallbacks": 0, "cached": 2}
        hg.graphs(False)
        assert hg.graph_stats()["cached"] == 0
        dims, rks = shapes[0]
        items = [([1.0], [O.rand_tt(dims, 4, np.random.default_rng(1))], None)]
        assert identical(rounding(T, hg, items, rks, orthogonal=False, seed=9)[0],
                         rounding(T, he, items, rks, orthogonal=False, seed=9)[0])
    finally:
        hg.close()
        he.close()
