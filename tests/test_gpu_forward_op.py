"""Forward (left-to-right) operator sketch and stta_sketch(x, L, R) on the device
against the oracle.

  tt_recursive_sketch(H, x, rks; reverse=false) (src/sketches.jl:390-459,
  contract_sketch_core_forwards! for operators :1240-1309): the device's W_k
  ((r_k, R_k, s_k), x rank fastest) against the oracle on the device's own
  blocks, 1e-12; the product H x is never formed.
  stta_sketch(x, L, R) (src/sketches.jl:1112-1124): Omega_k and Psi_k against
  the oracle's restatement, 1e-12.
  tt_sketch(...; reverse=false) and tt_combined_sketch(...; reverse=false)
  (src/sketches.jl:679-694,796-811,894-900,930-989): the right boundary sketch of
  a TT, of H x (never formed) and of a product (never formed), and the combined
  sketch with and without its seed + 1 extension, against the oracle on the
  device's blocks, 1e-12.
"""
import sys

import numpy as np
import pytest

import oracle as O
from helpers import to_api

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("orthogonal,block_rks", [(False, None), (True, 2)])
def test_forward_operator_sketch(orthogonal, block_rks):
    import ttround as T
    say(f"case: forward operator sketch orthogonal={orthogonal} block_rks={block_rks}")
    L = 5
    H = O.hubbard_1d_mpo(L, 1.0, 4.0)
    dims = H.dims
    x = O.rand_tt(dims, [1, 3, 5, 5, 3, 1], np.random.default_rng(21))
    rks = [1, 4, 6, 6, 4, 1]
    Hh = T.TToperator(H.N, [c.copy() for c in H.cores], H.dims, list(H.rks))
    W, srk = T.tt_recursive_sketch(Hh, to_api(x), rks, reverse=False, orthogonal=orthogonal, block_rks=block_rks,
                                   seed=91)
    blocks = {k: T.sketch_blocks(dims, rks, k, orthogonal=orthogonal, seed=91, block_rks=block_rks, forward=True)
              for k in range(L)}
    Wr, sr = O.tt_recursive_sketch_forward_op(H, x, rks, blocks.__getitem__, orthogonal, block_rks)
    assert list(srk) == list(sr)
    worst = 0.0
    for a, b in zip(W, Wr):
        assert a.shape == b.shape
        worst = max(worst, np.linalg.norm(a - b) / np.linalg.norm(b))
    say(f"  sketch ranks {srk}, worst rel err {worst:.3e}")
    assert worst < 1e-12


def test_stta_sketch_explicit_tts():
    import ttround as T
    say("case: stta_sketch(x, L, R) with explicit sketch TTs")
    dims = (3, 2, 4, 2, 3)
    g = np.random.default_rng(22)
    x = O.rand_tt(dims, [1, 3, 5, 6, 3, 1], g)
    Lt = O.rand_tt(dims, [1, 2, 4, 4, 3, 1], g)
    Rt = O.rand_tt(dims, [1, 3, 3, 4, 2, 1], g)
    Om, Ps = T.stta_sketch(to_api(x), to_api(Lt), to_api(Rt))
    Omr, Psr = O.stta_sketch_tt(x, Lt, Rt)
    for a, b in zip(Om, Omr):
        assert a.shape == b.shape and np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(b)
    for a, b in zip(Ps, Psr):
        assert a.shape == b.shape and np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(b)


def _fwd_provider(T, dims, rks, seed, orthogonal=False, p=None):
    cache = {}

    def get(k):
        if k not in cache:
            cache[k] = T.sketch_blocks(dims, rks, k, orthogonal=orthogonal, seed=seed, forward=True, p=p)
        return cache[k]
    return get


def test_forward_boundary_sketches():
    import ttround as T
    say("case: tt_sketch(...; reverse=false) of x, H x and (x, y)")
    L = 4
    H = O.hubbard_1d_mpo(L, 1.0, 4.0)
    dims, N = H.dims, L
    x = O.rand_tt(dims, [1, 3, 5, 4, 1], np.random.default_rng(23))
    y = O.rand_tt(dims, [1, 2, 3, 2, 1], np.random.default_rng(24))
    Hh = T.TToperator(H.N, [c.copy() for c in H.cores], H.dims, list(H.rks))
    s = 11
    bv, _ = O.sketch_structure_forward(dims, [1] * (N + 1))
    pc = -(-s // bv[N])
    pv = [1] + [pc] * N
    for orthogonal in (False, True):
        prov = _fwd_provider(T, dims, [1] * (N + 1), 33, orthogonal, pv)
        got, sa = T.tt_sketch(to_api(x), s, reverse=False, seed=33, orthogonal=orthogonal)
        ref, sr = O.tt_sketch(x, s, prov, orthogonal, reverse=False)
        assert sa == sr and len(got) == sa
        assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)
        got, sa = T.tt_sketch(Hh, to_api(x), s, reverse=False, seed=33, orthogonal=orthogonal)
        ref, sr = O.tt_sketch(x, s, prov, orthogonal, H=H, reverse=False)
        assert sa == sr and np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)
        got, sa = T.tt_sketch((to_api(x), to_api(y)), s, reverse=False, seed=33, orthogonal=orthogonal)
        ref, sr = O.tt_sketch_tuple([x, y], s, prov, orthogonal, reverse=False)
        assert sa == sr and np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)


@pytest.mark.parametrize("op", [False, True])
@pytest.mark.parametrize("s", [3, 25])
def test_forward_combined_sketch(op, s):
    """s = 3 fits the forward recursive sketch's right boundary rank; s = 25 needs the
    extra forward tt_sketch drawn from seed + 1."""
    import ttround as T
    say(f"case: tt_combined_sketch(...; reverse=false) op={op} s={s}")
    L = 4
    H = O.hubbard_1d_mpo(L, 1.0, 4.0)
    dims, N = H.dims, L
    x = O.rand_tt(dims, [1, 3, 5, 4, 1], np.random.default_rng(25))
    Hh = T.TToperator(H.N, [c.copy() for c in H.cores], H.dims, list(H.rks))
    rmax = 5
    args = ([Hh] if op else []) + [to_api(x), rmax, s]
    sk, W, srk = T.tt_combined_sketch(*args, reverse=False, orthogonal=False, seed=19)
    rks0 = [rmax] * (N + 1)
    rks0[0] = 1
    rks = list(rks0)
    rks[N] = max(rks[N], min(s, rks[N - 1]))
    prov = _fwd_provider(T, dims, rks, 19)
    bv, pr = O.sketch_structure_forward(dims, rks)
    bvs, _ = O.sketch_structure_forward(dims, [1] * (N + 1))
    pe = max(1, -(-(s - bv[N] * pr[N]) // bvs[N]))
    prov_e = _fwd_provider(T, dims, [1] * (N + 1), 20, p=[1] + [pe] * N)
    ref, Wr, sr = O.tt_combined_sketch(x, rks0, s, prov, prov_e, False, None, H=H if op else None, reverse=False)
    assert srk == sr and len(sk) == len(ref) and len(sk) >= s
    assert np.linalg.norm(sk - ref) <= 1e-12 * np.linalg.norm(ref)
    for a, b in zip(W, Wr):
        assert a.shape == b.shape
        assert np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(b)
