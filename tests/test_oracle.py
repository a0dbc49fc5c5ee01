"""Pins of the FP64 oracle (CPU only).

The reference ships no golden vectors for the rounding path (its tests are
tolerance invariants on unseeded random data), so every oracle function is
pinned here by something the mathematics fixes, chosen so that a dropped term,
a wrong index or a transposed operand fails:

  r_and_d_to_rks, block structure, rank heuristic  -> hand-worked integer cases
  floor_rule          -> closed-form singular-value vectors (tol 0, tails, degeneracy)
  tt_add, mpo_apply   -> dense brute force (test_tt_tools.jl:16-46 pattern)
  orthogonalize       -> dense reconstruction + per-core orthogonality (test_tt_tools.jl:48-57)
  tt_rounding         -> dense TT-SVD sweep on unfoldings (Oseledets 2011) + closed forms
  generate_sketch_blocks -> scaling constants, identity case, Q^T Q = c I and Q^T G upper triangular
  tt_recursive_sketch -> dense X_{k:N} Omega_{k:N} brute force; p=1 full sketch = dot
  tt_sketch           -> dense vec(x)^T Omega_{1:N} with constant p (TT and operator forms);
                         E|sketch|^2 = |x|^2 over seeds
  forward sketch      -> dense L_k^T Omega_{1:k-1} brute force
  stta_sketch         -> dense Omega_L^T X_(k) Omega_R and the three-index Psi_k; explicit-TT form
                         against dense interface matrices
  stable_inverse      -> A A^+ A = A, numpy pinv, the 1e-12 relative cutoff
  stta                -> exact recovery when the sketch ranks cover the ranks (test_ttrand.jl:136-147)
  tt_combined_sketch  -> parts equal the recursive sketch's W_1 and tt_sketch(seed + 1) with
                         the sqrt(len/s) weights; E|sketch|^2 = |x|^2 over seeds
  ttrand_rounding*    -> exact recovery when the sketch rank >= true rank (test_ttrand.jl:67-88)
  hubbard_1d_mpo      -> dense fermionic Hamiltonian
  hadamard            -> dense entrywise product (2 and 3 factors) and the bond order of each
                         layout; the factored NTuple sketch / rounding equal those of the
                         explicit product, which the rounding recovers exactly
  dot, dot_par, partial_contraction, dot_operator
                      -> dense inner products, dense left/right interface
                         contractions, psi^T H phi with the dense operator
  Philox              -> Random123 known-answer vectors
"""
import math

import numpy as np
import pytest

import oracle as O
from oracle import philox as P

RNG = np.random.default_rng(20240601)


def rand_tt(dims, rks, seed):
    return O.rand_tt(dims, rks, np.random.default_rng(seed))


# ---------------------------------------------------------------- integers

def test_r_and_d_to_rks_hand_cases():
    assert O.r_and_d_to_rks([1, 100, 100, 100, 100, 1], (4, 4, 4, 4, 4), rmax=100) == [1, 4, 16, 16, 4, 1]
    assert O.r_and_d_to_rks([1, 5, 5, 1], (2, 3, 4)) == [1, 2, 4, 1]
    # d=40, n=4: prod(dims[1:i-1]) overflows Int64 to 0 for i >= 33 and is then ignored
    r = O.r_and_d_to_rks([256] * 41, (4,) * 40, rmax=256)
    assert r[:5] == [1, 4, 16, 64, 256] and r[-5:] == [256, 64, 16, 4, 1]
    assert O.tt_oracle._wrapped_prod([4] * 32) == 0 and O.tt_oracle._wrapped_prod([2] * 63) == -(1 << 63)


def test_block_structure_c1():
    # configs C1 with block_rks = 144: s = [144, ..., 144, 192, 144, 144, 1]
    d = 32
    rks = [1] + [144] * (d - 1) + [1]
    bv, p = O.sketch_structure((4,) * d, rks, 144)
    s = [b * q for b, q in zip(bv, p)]
    assert s[1:29] == [144] * 28 and s[29:] == [192, 144, 144, 1]
    assert bv[-5:] == [144, 64, 16, 4, 1] and p[-5:] == [1, 3, 9, 36, 1]


def test_block_structure_default_and_monotone():
    dims = (2,) * 6
    bv, p = O.sketch_structure(dims, [1, 2, 4, 6, 4, 2, 1], None)
    assert bv == [6, 6, 6, 6, 4, 2, 1]
    assert p == [1, 1, 1, 1, 1, 1, 1]
    bv, p = O.sketch_structure(dims, [1, 9, 9, 9, 9, 9, 1], 2)
    assert bv == [2, 2, 2, 2, 2, 2, 1] and p == [1, 5, 5, 5, 5, 5, 1]
    bv, p = O.sketch_structure(dims, [1, 9, 9, 9, 9, 9, 1], 2, oversampling=1.5)
    assert p == [2, 8, 8, 8, 8, 8, 1]


def test_default_rank_heuristic():
    a = O.zeros_tt((3, 3, 3, 3), [1, 3, 5, 4, 1])
    b = O.zeros_tt((3, 3, 3, 3), [1, 4, 6, 3, 1])
    assert O.default_rank_heuristic(a) == [1, 6, 8, 7, 1]
    assert O.default_rank_heuristic([a, b]) == [1, 7, 9, 7, 1]


# ---------------------------------------------------------------- floor

def test_floor_rule_closed_forms():
    s = np.array([4.0, 3.0, 2.0, 1.0])
    assert O.floor_rule(s, 0.0)[1] == 4
    # tail energy: drop s4 (1 < tol^2*30 = 1.2) but not s3 (4+1 = 5)
    assert O.floor_rule(s, 0.2)[1] == 3
    assert O.floor_rule(s, math.sqrt(5.0 / 30.0) + 1e-12)[1] == 2
    assert O.floor_rule(s, 1e-12)[1] == 4
    # degeneracy extension with atol 1e-5 on s[k+1]/s[k]
    # (tol^2 = 1.44e-6 lies between s3^2 + s4^2 and s2^2 + s3^2 + s4^2, so the cut is k = 2)
    s = np.array([1.0, 1e-3, 1e-3 * (1 - 5e-6), 1e-9])
    assert O.floor_rule(s, 1.2e-3)[1] == 3
    s = np.array([1.0, 1e-3, 1e-3 * (1 - 2e-5), 1e-9])
    assert O.floor_rule(s, 1.2e-3)[1] == 2
    s = np.array([1.0, 0.0, 0.0])
    assert O.floor_rule(s, 1e-3)[1] == 1


# ---------------------------------------------------------------- algebra

def test_add_scale_lincomb_dense():
    dims = (3, 2, 4, 3)
    x, y, z = rand_tt(dims, 3, 1), rand_tt(dims, 2, 2), rand_tt(dims, 4, 3)
    X, Y, Z = (O.ttv_to_tensor(t) for t in (x, y, z))
    np.testing.assert_allclose(O.ttv_to_tensor(O.tt_add(x, y)), X + Y, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(O.ttv_to_tensor(O.tt_sub(x, y)), X - Y, rtol=1e-13, atol=1e-13)
    lc = O.tt_lincomb([x, y, z], [0.5, -2.0, 3.0])
    np.testing.assert_allclose(O.ttv_to_tensor(lc), 0.5 * X - 2 * Y + 3 * Z, rtol=1e-12, atol=1e-12)
    assert lc.rks == [1, 3 + 2 + 3, 9, 9, 1] or lc.rks[1] == x.rks[1] + y.rks[1] + z.rks[1]
    assert O.tt_scale(0.0, x).rks == [1] * 5


def test_mpo_apply_dense():
    dims = (2, 3, 2)
    A = O.TTO([RNG.standard_normal((n, n, ra, rb)) for n, ra, rb in zip(dims, [1, 2, 3], [2, 3, 1])], dims, [1, 2, 3, 1])
    x = rand_tt(dims, [1, 2, 2, 1], 4)
    Ad = O.tto_to_tensor(A)
    xd = O.ttv_to_tensor(x).reshape(-1, order="F")
    y = O.mpo_apply(A, x)
    assert y.rks == [1, 4, 6, 1]
    np.testing.assert_allclose(O.ttv_to_tensor(y).reshape(-1, order="F"), Ad @ xd, rtol=1e-12, atol=1e-12)


def test_hubbard_mpo_dense():
    for L in (2, 3, 4):
        H = O.hubbard_1d_mpo(L, 1.0, 4.0)
        assert H.rks == [1] + [6] * (L - 1) + [1]
        np.testing.assert_allclose(O.tto_to_tensor(H), O.hubbard_1d_dense(L, 1.0, 4.0), atol=1e-13)


# ---------------------------------------------------------------- inner products

def _left_part(x, k):
    """Dense (n_1...n_k, r_k) left interface matrix of cores 1..k."""
    acc = np.ones((1, 1))
    for j in range(k):
        c = x.cores[j]
        acc = np.einsum("ma,iab->mib", acc, c).reshape(-1, c.shape[2], order="F")
    return acc


def _right_part(x, k):
    """Dense (r_k, n_{k+1}...n_N) right interface matrix of cores k+1..N (0-based k)."""
    acc = np.ones((1, 1))
    for j in range(x.N - 1, k - 1, -1):
        c = x.cores[j]
        acc = np.einsum("iab,bm->aim", c, acc).reshape(c.shape[1], -1, order="F")
    return acc


def test_dot_and_dot_par_dense():
    dims = (3, 2, 4, 2, 3)
    x, y = rand_tt(dims, [1, 3, 5, 4, 2, 1], 21), rand_tt(dims, [1, 2, 6, 3, 3, 1], 22)
    ref = float(np.vdot(O.ttv_to_tensor(x), O.ttv_to_tensor(y)))
    for f in (O.dot, O.dot_par, O.tt_dot):
        assert abs(f(x, y) - ref) <= 1e-12 * abs(ref)
    x1, y1 = rand_tt((7,), [1, 1], 23), rand_tt((7,), [1, 1], 24)
    assert abs(O.dot(x1, y1) - float(x1.cores[0].ravel() @ y1.cores[0].ravel())) < 1e-13


@pytest.mark.parametrize("reverse", [True, False])
def test_partial_contraction_dense(reverse):
    dims = (2, 3, 2, 4)
    x, y = rand_tt(dims, [1, 2, 5, 3, 1], 31), rand_tt(dims, [1, 3, 4, 2, 1], 32)
    W = O.partial_contraction(x, y, reverse=reverse)
    assert [w.shape for w in W] == [(x.rks[k], y.rks[k]) for k in range(5)]
    for k in range(5):
        if reverse:
            ref = _right_part(x, k) @ _right_part(y, k).T
        else:
            ref = _left_part(x, k).T @ _left_part(y, k)
        np.testing.assert_allclose(W[k], ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("ots", [(None, None), ((1, 0, -1, -1), None), (None, (0, -1, -1, -1)),
                                 ((1, 1, 1, 0), (1, 0, -1, -1)), ((1, 1, 0, -1), (1, 1, 1, 0))])
def test_dot_operator_dense(ots):
    dims = (2, 3, 2, 2)
    psi, phi = rand_tt(dims, [1, 2, 4, 3, 1], 41), rand_tt(dims, [1, 3, 3, 2, 1], 42)
    # the pivot only changes the contraction order, so any ot labels are valid here
    if ots[0] is not None:
        psi.ot = list(ots[0])
    if ots[1] is not None:
        phi.ot = list(ots[1])
    Hr = [1, 3, 2, 4, 1]
    H = O.TTO([RNG.standard_normal((n, n, Hr[k], Hr[k + 1])) for k, n in enumerate(dims)], dims, Hr)
    ref = float(O.ttv_to_tensor(psi).reshape(-1, order="F") @ O.tto_to_tensor(H)
                @ O.ttv_to_tensor(phi).reshape(-1, order="F"))
    assert abs(O.dot_operator(psi, H, phi) - ref) <= 1e-12 * abs(ref)
    # Hubbard energy of a state against the dense Hamiltonian
    Hh = O.hubbard_1d_mpo(3, 1.0, 4.0)
    v = rand_tt((4, 4, 4), [1, 4, 4, 1], 43)
    vd = O.ttv_to_tensor(v).reshape(-1, order="F")
    e = O.dot_operator(v, Hh, v)
    assert abs(e - vd @ O.hubbard_1d_dense(3, 1.0, 4.0) @ vd) <= 1e-12 * abs(e)


# ---------------------------------------------------------------- deterministic rounding

@pytest.mark.parametrize("root", [1, 3, 6])
def test_orthogonalize(root):
    dims = (2, 3, 2, 2, 3, 2)
    x = rand_tt(dims, [1, 2, 5, 7, 6, 4, 1], 11)
    y = O.orthogonalize(x, i=root)
    np.testing.assert_allclose(O.ttv_to_tensor(y), O.ttv_to_tensor(x), rtol=1e-12, atol=1e-12)
    for k in range(x.N):
        c = y.cores[k]
        if k < root - 1:  # left-orthogonal: sum_i A_i^T A_i = I
            G = sum(c[i].T @ c[i] for i in range(c.shape[0]))
            np.testing.assert_allclose(G, np.eye(c.shape[2]), atol=1e-13)
            assert y.ot[k] == 1
        elif k > root - 1:  # right-orthogonal: sum_i A_i A_i^T = I
            G = sum(c[i] @ c[i].T for i in range(c.shape[0]))
            np.testing.assert_allclose(G, np.eye(c.shape[1]), atol=1e-13)
            assert y.ot[k] == -1
        else:
            assert y.ot[k] == 0
    assert abs(O.tt_norm(x) - np.linalg.norm(O.ttv_to_tensor(x))) < 1e-12 * np.linalg.norm(O.ttv_to_tensor(x))


def dense_tt_svd_left(T, dims, tol, rmax):
    """Right-to-left truncation of the dense unfoldings with the floor rule."""
    T = T.copy()
    d = len(dims)
    ranks = [1] * (d + 1)
    for j in range(d, 1, -1):  # bond between modes j-1 | j (1-based)
        rows = int(np.prod(dims[: j - 1]))
        M = np.reshape(T, (rows, -1), order="F")
        u, s, vt = np.linalg.svd(M, full_matrices=False)
        _, k = O.floor_rule(s, tol)
        k = min(k, rmax, int(np.sum(s > 0)))
        M = (u[:, :k] * s[:k]) @ vt[:k]
        T = np.reshape(M, T.shape, order="F")
        ranks[j - 1] = k
    return T, ranks


@pytest.mark.parametrize("tol,rmax", [(1e-12, 2 ** 14), (1e-2, 2 ** 14), (1e-12, 3), (0.3, 5)])
def test_tt_rounding_matches_dense_svd(tol, rmax):
    dims = (2, 3, 2, 3, 2)
    x = O.tt_add(rand_tt(dims, 3, 5), O.tt_scale(0.05, rand_tt(dims, 4, 6)))
    X = O.ttv_to_tensor(x)
    Td, ranks = dense_tt_svd_left(X, dims, tol, rmax)
    y = O.tt_rounding(x, tol=tol, rmax=rmax)
    assert y.rks[1:-1] == ranks[1:-1]
    np.testing.assert_allclose(O.ttv_to_tensor(y), Td, rtol=1e-11, atol=1e-11 * np.abs(X).max())
    assert y.ot == [0] + [-1] * (len(dims) - 1)
    yr = O.tt_rounding(x, tol=tol, rmax=rmax, direction="right")
    assert yr.ot == [1] * (len(dims) - 1) + [0]


def test_tt_rounding_closed_form():
    # C0 pattern: round(y + y + 1e-10 z; 1e-8) = 2y with the ranks of y
    dims = (4,) * 6
    y, z = rand_tt(dims, 5, 7), rand_tt(dims, 5, 8)
    x = O.tt_add(O.tt_add(y, y), O.tt_scale(1e-10, z))
    r = O.tt_rounding(x, tol=1e-8)
    assert r.rks == y.rks
    assert O.rel_err(O.tt_scale(2.0, y), r) < 1e-9
    with pytest.raises(ValueError):
        O.tt_rounding(x, direction="up")


# ---------------------------------------------------------------- sketches

def test_generate_sketch_blocks():
    g = RNG.standard_normal(3 * 2 * 4 * 2)
    S = O.generate_sketch_blocks(g, 3, 2, 4, 2, orthogonal=False)
    np.testing.assert_allclose(O.jreshape(S, -1), g / 2.0, rtol=1e-15)
    S = O.generate_sketch_blocks(g, 3, 2, 4, 2, orthogonal=True)
    G = O.jreshape(g, (6, 4, 2))
    for j in range(2):
        Q = O.jreshape(S[:, :, :, j], (6, 4))
        np.testing.assert_allclose(Q.T @ Q, (6 / 4) * np.eye(4), atol=1e-13)
        R = Q.T @ G[:, :, j] / math.sqrt(6 / 4)
        np.testing.assert_allclose(np.tril(R, -1), 0.0, atol=1e-12)
    S = O.generate_sketch_blocks(None, 2, 3, 6, 2, orthogonal=True)
    for j in range(2):
        np.testing.assert_array_equal(O.jreshape(S[:, :, :, j], (6, 6)), np.eye(6))


def _dense_right(x, k):
    """X_{k:N} as an (r_k x prod n) matrix (column-major multi-index)."""
    acc = x.cores[k].transpose(1, 0, 2)  # (r, n, r')
    for j in range(k + 1, x.N):
        acc = np.tensordot(acc, x.cores[j], axes=([acc.ndim - 1], [1]))
    return acc[..., 0].reshape(x.rks[k], -1, order="F")


def _dense_sketch(Sblocks, dims, bv, p, k):
    """Omega_{k:N} as (prod n) x (bv_k p_k) from the block chains."""
    N = len(dims)
    cols = []
    for j in range(p[k]):
        acc = Sblocks[k][:, :, :, j]  # (bvr, n, bvl)
        acc = acc.transpose(1, 2, 0)  # (n, bvl, bvr)
        # build Omega^{(j)}_{k:N}[(i_k..i_N), b]
        om = None
        for q in range(N - 1, k - 1, -1):
            jj = j if q < N - 1 else 0
            Sq = Sblocks[q][:, :, :, j if q > k else j]  # (bvr, n, bvl)
            if q == N - 1:
                om = Sq[0]  # (n, bvl)
            else:
                # om_new[(i_q, rest), b] = sum_beta S_q[beta, i_q, b] om[rest, beta]
                t = np.einsum("bic,rb->irc", Sq, om)
                om = t.reshape(-1, Sq.shape[2], order="F")
            del jj
        cols.append(om)
    return np.concatenate(cols, axis=1)


@pytest.mark.parametrize("orthogonal,block_rks", [(False, None), (True, None), (False, 2), (True, 3)])
def test_recursive_sketch_dense(orthogonal, block_rks):
    dims = (2, 3, 2, 2, 3)
    N = len(dims)
    x = rand_tt(dims, [1, 2, 4, 5, 3, 1], 21)
    rks = [1, 4, 5, 5, 4, 1]
    bv, p = O.sketch_structure(dims, rks, block_rks)
    draws = [RNG.standard_normal(bv[k + 1] * dims[k] * bv[k] * p[k]) for k in range(N)]
    W, s = O.tt_recursive_sketch(x, rks, draws, orthogonal, block_rks)
    Sb = [O.generate_sketch_blocks(draws[k], bv[k + 1], dims[k], bv[k], p[k], orthogonal) for k in range(N)]
    for k in range(N):
        Om = _dense_sketch(Sb, dims, bv, p, k)
        ref = _dense_right(x, k) @ Om / math.sqrt(p[k])
        np.testing.assert_allclose(W[k], ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
        assert W[k].shape == (x.rks[k], s[k])


def test_full_sketch_with_p1_is_dot():
    # with block rank >= every bond the sketch is a Gaussian TT and W[1] = <x, omega>
    dims = (2, 2, 2, 2)
    x = rand_tt(dims, [1, 2, 4, 2, 1], 3)
    rks = [1, 1, 1, 1, 1]
    bv, p = O.sketch_structure(dims, rks, 1)
    draws = [RNG.standard_normal(bv[k + 1] * dims[k] * bv[k] * p[k]) for k in range(4)]
    W, _ = O.tt_recursive_sketch(x, rks, draws, False, 1)
    om = O.TT([O.generate_sketch_blocks(draws[k], 1, 2, 1, 1, False)[:, :, :, 0].transpose(1, 0, 2) for k in range(4)],
              dims, [1] * 5)
    assert abs(W[0][0, 0] - O.tt_dot(x, om)) < 1e-12 * abs(O.tt_dot(x, om)) + 1e-14


def _rand_tto(dims, rks, seed):
    g = np.random.default_rng(seed)
    return O.TTO([g.standard_normal((n, n, rks[k], rks[k + 1])) for k, n in enumerate(dims)], dims, rks)


@pytest.mark.parametrize("orthogonal,block_rks,op", [(False, None, False), (True, None, False), (False, 2, True),
                                                     (True, 3, True)])
def test_tt_sketch_dense(orthogonal, block_rks, op):
    # tt_sketch (sketches.jl:645-813): vec(x)^T Omega_{1:N} / sqrt(p), the dense Omega
    # built from the block chains with the same p at every core
    dims = (2, 3, 2, 2, 3)
    N = len(dims)
    x = rand_tt(dims, [1, 2, 4, 5, 3, 1], 23)
    H = _rand_tto(dims, [1, 2, 3, 2, 2, 1], 24) if op else None
    target = O.mpo_apply(H, x) if op else x
    s = 7
    bv, _ = O.sketch_structure(dims, [1] * (N + 1), block_rks)
    p = -(-s // bv[0])
    draws = [RNG.standard_normal(bv[k + 1] * dims[k] * bv[k] * p) for k in range(N)]
    got, sa = O.tt_sketch(x, s, draws, orthogonal, block_rks, H=H)
    assert sa == bv[0] * p and got.shape == (sa,) and sa >= s
    Sb = [O.generate_sketch_blocks(draws[k], bv[k + 1], dims[k], bv[k], p, orthogonal) for k in range(N)]
    Om = _dense_sketch(Sb, dims, bv, [p] * N + [1], 0)
    ref = (_dense_right(target, 0) @ Om)[0] / math.sqrt(p)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_tt_sketch_and_combined_sketch_are_unbiased():
    # Gaussian blocks scaled by 1/sqrt(bv_k), boundary divided by sqrt(p):
    # E |sketch|^2 = |x|^2; the combined sketch keeps that only with the
    # sqrt(len/s) weights of sketches.jl:947-948
    dims = (2, 2, 3, 2)
    N = len(dims)
    x = rand_tt(dims, [1, 2, 3, 2, 1], 25)
    nx2 = O.tt_norm(x) ** 2
    g = np.random.default_rng(99)
    rks = [1, 3, 3, 3, 1]
    acc_s, acc_c, trials = 0.0, 0.0, 600
    bvr, pr = O.sketch_structure(dims, [max(rks[0], min(9, rks[1]))] + rks[1:], None)
    bvs, _ = O.sketch_structure(dims, [1] * (N + 1), None)
    for _ in range(trials):
        dr = [g.standard_normal(bvr[k + 1] * dims[k] * bvr[k] * pr[k]) for k in range(N)]
        de = [g.standard_normal(bvs[k + 1] * dims[k] * bvs[k] * 4) for k in range(N)]
        sk, _ = O.tt_sketch(x, 5, de, False)
        acc_s += float(sk @ sk)
        cs, W, srk = O.tt_combined_sketch(x, rks, 9, dr, de, False)
        assert len(cs) >= 9 and srk[0] == 1 and W[0].shape == (1, 1)
        acc_c += float(cs @ cs)
    assert abs(acc_s / trials / nx2 - 1.0) < 0.15
    assert abs(acc_c / trials / nx2 - 1.0) < 0.15


def test_tt_combined_sketch_parts():
    dims = (2, 3, 2, 2)
    N = len(dims)
    x = rand_tt(dims, [1, 2, 4, 3, 1], 26)
    rks = [1, 3, 4, 3, 1]
    # s within the recursive sketch's boundary rank: no extra sketch
    rk1 = [max(1, min(2, rks[1]))] + rks[1:]
    bv, p = O.sketch_structure(dims, rk1, None)
    dr = [RNG.standard_normal(bv[k + 1] * dims[k] * bv[k] * p[k]) for k in range(N)]
    cs, W, srk = O.tt_combined_sketch(x, rks, 2, dr, None, False)
    Wr, sr = O.tt_recursive_sketch(x, rk1, dr, False)
    np.testing.assert_array_equal(cs, O.jreshape(Wr[0], -1))
    assert srk == [1] + sr[1:] and all(np.array_equal(a, b) for a, b in zip(W[1:], Wr[1:]))
    # s above it: [sqrt(l/s) W_1 ; sqrt(le/s) tt_sketch(seed + 1)]
    s = 11
    rk2 = [max(1, min(s, rks[1]))] + rks[1:]
    bv, p = O.sketch_structure(dims, rk2, None)
    dr = [RNG.standard_normal(bv[k + 1] * dims[k] * bv[k] * p[k]) for k in range(N)]
    bvs, _ = O.sketch_structure(dims, [1] * (N + 1), None)
    Wr, sr = O.tt_recursive_sketch(x, rk2, dr, False)
    pe = -(-(s - sr[0]) // bvs[0])
    de = [RNG.standard_normal(bvs[k + 1] * dims[k] * bvs[k] * pe) for k in range(N)]
    ex, le = O.tt_sketch(x, s - sr[0], de, False)
    cs, W, srk = O.tt_combined_sketch(x, rks, s, dr, de, False)
    tot = sr[0] + le
    assert len(cs) == tot and tot >= s
    np.testing.assert_allclose(cs[: sr[0]], math.sqrt(sr[0] / tot) * O.jreshape(Wr[0], -1), rtol=1e-15)
    np.testing.assert_allclose(cs[sr[0]:], math.sqrt(le / tot) * ex, rtol=1e-15)


def _dense_sketch_fwd(Sblocks, dims, p, k):
    """Forward Omega_{1:k} as (n_1...n_k) x (bv_{k+1} p_{k+1}) from the block chains."""
    cols = []
    for j in range(p[k]):
        om = Sblocks[0][0, :, :, j]  # (n_1, bv_2)
        for q in range(1, k):
            t = np.einsum("mb,bic->mic", om, Sblocks[q][:, :, :, j])
            om = t.reshape(-1, Sblocks[q].shape[2], order="F")
        cols.append(om)
    return np.concatenate(cols, axis=1)


@pytest.mark.parametrize("orthogonal,block_rks", [(False, None), (True, None), (False, 2), (True, 3)])
def test_forward_sketch_dense(orthogonal, block_rks):
    # tt_recursive_sketch(...; reverse=false) (sketches.jl:241-303): W_k = L_k^T Omega_{1:k-1} / sqrt(p_k)
    dims = (2, 3, 2, 2, 3)
    N = len(dims)
    x = rand_tt(dims, [1, 2, 4, 5, 3, 1], 27)
    rks = [1, 4, 5, 5, 4, 1]
    bv, p = O.sketch_structure_forward(dims, rks, block_rks, oversampling=1.5)
    assert p[0] == 1 and all(p[k] >= p[k + 1] for k in range(1, N))
    draws = [RNG.standard_normal(bv[k] * dims[k] * bv[k + 1] * p[k + 1]) for k in range(N)]
    W, s = O.tt_recursive_sketch_forward(x, rks, draws, orthogonal, block_rks, oversampling=1.5)
    Sb = [O.generate_sketch_blocks(draws[k], bv[k], dims[k], bv[k + 1], p[k + 1], orthogonal) for k in range(N)]
    for k in range(1, N + 1):
        ref = _left_part(x, k).T @ _dense_sketch_fwd(Sb, dims, p, k) / math.sqrt(p[k])
        np.testing.assert_allclose(W[k], ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
        assert W[k].shape == (x.rks[k], s[k])


def _stta_draws(dims, rks, block_rks=None, g=None):
    g = g or RNG
    bvl, pl = O.sketch_structure_forward(dims, rks, block_rks, oversampling=1.5)
    bvr, pr = O.sketch_structure(dims, rks, block_rks)
    N = len(dims)
    dl = [g.standard_normal(bvl[k] * dims[k] * bvl[k + 1] * pl[k + 1]) for k in range(N)]
    dr = [g.standard_normal(bvr[k + 1] * dims[k] * bvr[k] * pr[k]) for k in range(N)]
    return dl, dr, (bvl, pl), (bvr, pr)


@pytest.mark.parametrize("orthogonal", [False, True])
def test_stta_sketch_dense(orthogonal):
    # Omega_k = Omega_L^T X_(k) Omega_R, Psi_k = Omega_L^T X Omega_R over the outer modes
    dims = (2, 3, 2, 3)
    N = len(dims)
    x = rand_tt(dims, [1, 2, 3, 2, 1], 28)
    rks = [1, 2, 3, 2, 1]
    dl, dr, (bvl, pl), (bvr, pr) = _stta_draws(dims, rks)
    Om, Ps, sl, sr = O.stta_sketch(x, rks, dl, dr, orthogonal)
    SL = [O.generate_sketch_blocks(dl[k], bvl[k], dims[k], bvl[k + 1], pl[k + 1], orthogonal) for k in range(N)]
    SR = [O.generate_sketch_blocks(dr[k], bvr[k + 1], dims[k], bvr[k], pr[k], orthogonal) for k in range(N)]
    X = O.ttv_to_tensor(x)
    for k in range(N - 1):
        nl = int(np.prod(dims[: k + 1]))
        XL = X.reshape(nl, -1, order="F")
        OL = _dense_sketch_fwd(SL, dims, pl, k + 1) / math.sqrt(pl[k + 1])
        OR = _dense_sketch(SR, dims, bvr, pr, k + 1) / math.sqrt(pr[k + 1])
        ref = OL.T @ XL @ OR
        assert Om[k].shape == (sl[k + 1], sr[k + 1])
        np.testing.assert_allclose(Om[k], ref, rtol=1e-11, atol=1e-11 * np.abs(ref).max())
    for k in range(N):
        m0 = int(np.prod(dims[:k]))
        X3 = X.reshape(m0, dims[k], -1, order="F")
        OL = (_dense_sketch_fwd(SL, dims, pl, k) / math.sqrt(pl[k])) if k > 0 else np.ones((1, 1))
        OR = (_dense_sketch(SR, dims, bvr, pr, k + 1) / math.sqrt(pr[k + 1])) if k < N - 1 else np.ones((1, 1))
        ref = np.einsum("ma,mib,bc->iac", OL, X3, OR, optimize=True)
        assert Ps[k].shape == (dims[k], sl[k], sr[k + 1])
        np.testing.assert_allclose(Ps[k], ref, rtol=1e-11, atol=1e-11 * np.abs(ref).max())


def test_stta_sketch_with_explicit_tts_dense():
    # stta_sketch(x, L, R) (sketches.jl:1112-1124): Omega_k = L_{<=k}^T X_(k) R_{>k}^T
    dims = (2, 3, 2, 2)
    N = len(dims)
    x = rand_tt(dims, [1, 2, 4, 2, 1], 29)
    L = rand_tt(dims, [1, 2, 5, 3, 1], 30)
    R = rand_tt(dims, [1, 2, 3, 2, 1], 31)
    Om, Ps = O.stta_sketch_tt(x, L, R)
    X = O.ttv_to_tensor(x)
    for k in range(N - 1):
        XL = X.reshape(int(np.prod(dims[: k + 1])), -1, order="F")
        ref = _left_part(L, k + 1).T @ XL @ _right_part(R, k + 1).T
        np.testing.assert_allclose(Om[k], ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    for k in range(N):
        X3 = X.reshape(int(np.prod(dims[:k])), dims[k], -1, order="F")
        ref = np.einsum("ma,mib,cb->iac", _left_part(L, k), X3, _right_part(R, k + 1), optimize=True)
        np.testing.assert_allclose(Ps[k], ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_stable_inverse():
    A = RNG.standard_normal((7, 4))
    P = O.stable_inverse(A)
    np.testing.assert_allclose(A @ P @ A, A, atol=1e-12)
    np.testing.assert_allclose(P, np.linalg.pinv(A), atol=1e-12)
    # singular values below 1e-12 max(s) are dropped (src/tt_randtools.jl:485-488)
    u, _, vt = np.linalg.svd(RNG.standard_normal((6, 6)))
    B = u @ np.diag([3.0, 2.0, 1.0, 0.5, 1e-13, 1e-14]) @ vt
    np.testing.assert_allclose(O.stable_inverse(B), vt[:4].T @ np.diag(1 / np.array([3.0, 2.0, 1.0, 0.5])) @ u[:, :4].T,
                               atol=1e-10)


@pytest.mark.parametrize("orthogonal", [False, True])
def test_stta_exact_when_rank_sufficient(orthogonal):
    # test_ttrand.jl:136-147 pattern: STTA reproduces x when the sketch ranks cover its ranks
    dims = (2,) * 7
    x = rand_tt(dims, [1, 2, 4, 4, 4, 4, 2, 1], 32)
    rks = [1, 2, 4, 6, 6, 4, 2, 1]
    dl, dr, _, _ = _stta_draws(dims, rks)
    y = O.stta([1.0], [x], rks, dl, dr, orthogonal)
    assert O.rel_err(x, y) < 1e-10
    ys = [rand_tt(dims, [1, 2, 2, 2, 2, 2, 2, 1], 33 + j) for j in range(2)]
    y2 = O.stta([0.5, -2.0], ys, rks, dl, dr, orthogonal)
    assert O.rel_err(O.tt_lincomb(ys, [0.5, -2.0]), y2) < 1e-10


# ---------------------------------------------------------------- randomized rounding

def _draws_for(dims, rks, block_rks, seed=1234):
    bv, p = O.sketch_structure(dims, rks, block_rks)
    return [P.normal_draws(seed, k, bv[k + 1] * dims[k] * bv[k] * p[k]) for k in range(len(dims))]


@pytest.mark.parametrize("orthogonal", [False, True])
def test_ttrand_rounding_exact_when_rank_sufficient(orthogonal):
    dims = (2,) * 8
    a = rand_tt(dims, [1, 2, 4, 4, 4, 4, 4, 2, 1], 31)
    pert = O.tt_add(a, O.tt_scale(1e-8, rand_tt(dims, [1, 2, 3, 3, 3, 3, 3, 2, 1], 32)))
    rks = [1, 2, 4, 8, 16, 8, 4, 2, 1]
    y = O.ttrand_rounding(pert, rks, _draws_for(dims, rks, None), orthogonal)
    assert O.rel_err(pert, y) < 1e-12
    for k in range(len(dims) - 1):
        c = y.cores[k]
        G = sum(c[i].T @ c[i] for i in range(c.shape[0]))
        np.testing.assert_allclose(G, np.eye(c.shape[2]), atol=1e-12)
    assert y.ot == [1] * 7 + [0]


def test_ttrand_rounding_lincomb_and_sum_agree_exactly_recovered():
    dims = (2,) * 6
    ys = [rand_tt(dims, 2, 40 + j) for j in range(3)]
    al = [0.5, -1.0, 2.0]
    rks = [1, 2, 4, 8, 4, 2, 1]
    dr = _draws_for(dims, rks, None)
    y1 = O.ttrand_rounding_lincomb(al, ys, rks, dr, False)
    ref = O.tt_lincomb(ys, al)
    assert O.rel_err(ref, y1) < 1e-12


def test_ttrand_rounding_truncates_noise():
    # test_ttrand.jl:164-175 pattern: |A - rand(A + eps B, r)| < 1e-3 |A|
    dims = (6,) * 6
    A = rand_tt(dims, 3, 50)
    B = rand_tt(dims, 12, 51)
    eps = 1e-4
    x = O.tt_add(O.tt_scale(1.0 / O.tt_norm(A), A), O.tt_scale(eps / O.tt_norm(B), B))
    rks = [1] + [6] * 5 + [1]
    y = O.ttrand_rounding(x, rks, _draws_for(dims, rks, None), True)
    assert O.rel_err(O.tt_scale(1.0 / O.tt_norm(A), A), y) < 1e-3


def test_ttrand_rounding_operator_exact():
    L = 4
    H = O.hubbard_1d_mpo(L, 1.0, 4.0)
    dims = H.dims
    y = rand_tt(dims, [1, 3, 5, 3, 1], 60)
    Hy = O.mpo_apply(H, y)
    rks = [1, 4, 16, 4, 1]
    dr = _draws_for(dims, rks, None)
    r = O.ttrand_rounding_op([1.5], H, [y], rks, dr, False)
    assert O.rel_err(O.tt_scale(1.5, Hy), r) < 1e-11
    # operator sketch equals the plain sketch of the explicit product up to the bond order
    Wop, s1 = O.tt_recursive_sketch_op(H, y, rks, dr, False)
    Wex, s2 = O.tt_recursive_sketch(Hy, rks, dr, False)
    assert s1 == s2
    for k in range(L + 1):
        ex = Wex[k].reshape(H.rks[k], y.rks[k], -1, order="F").transpose(1, 0, 2)
        np.testing.assert_allclose(Wop[k], ex, rtol=1e-11, atol=1e-11 * np.abs(ex).max())
    b = rand_tt(dims, [1, 2, 2, 2, 1], 61)
    res = O.ttrand_rounding_residual(H, y, b, rks, dr, False)
    assert O.rel_err(O.tt_sub(Hy, b), res) < 1e-11


# ---------------------------------------------------------------- entrywise products

@pytest.mark.parametrize("first_fastest", [False, True])
def test_hadamard_dense(first_fastest):
    dims = (2, 3, 2, 3)
    x = rand_tt(dims, [1, 2, 3, 2, 1], 70)
    y = rand_tt(dims, [1, 3, 2, 4, 1], 71)
    z = rand_tt(dims, [1, 2, 2, 2, 1], 72)
    p2 = O.hadamard([x, y], first_fastest)
    np.testing.assert_allclose(O.ttv_to_tensor(p2), O.ttv_to_tensor(x) * O.ttv_to_tensor(y), rtol=1e-12, atol=1e-13)
    assert p2.rks == [a * b for a, b in zip(x.rks, y.rks)] and p2.ot == [0, 0, 0, 0]
    p3 = O.hadamard([x, y, z], first_fastest)
    np.testing.assert_allclose(O.ttv_to_tensor(p3), O.ttv_to_tensor(x) * O.ttv_to_tensor(y) * O.ttv_to_tensor(z),
                               rtol=1e-12, atol=1e-13)
    # bond order: kron(x, y) puts y's rank index fastest (src/tt_operations.jl:283);
    # the tuple order puts x's fastest
    rxa, rxb, rya, ryb = x.rks[1], x.rks[2], y.rks[1], y.rks[2]
    for a in range(rxa * rya):
        for b in range(rxb * ryb):
            if first_fastest:
                da, db = (a % rxa, a // rxa), (b % rxb, b // rxb)
            else:
                da, db = (a // rya, a % rya), (b // ryb, b % ryb)
            ref = x.cores[1][1, da[0], db[0]] * y.cores[1][1, da[1], db[1]]
            assert p2.cores[1][1, a, b] == ref


def test_hadamard_sketch_and_rounding_factored():
    """The factored NTuple sketch and rounding (product never formed) equal the
    plain sketch and rounding of the explicit product in tuple order; the
    rounding recovers the product exactly when the ranks cover it."""
    dims = (2, 3, 2, 2)
    N = len(dims)
    x = rand_tt(dims, [1, 2, 3, 2, 1], 73)
    y = rand_tt(dims, [1, 2, 2, 3, 1], 74)
    xy = O.hadamard([x, y], first_fastest=True)
    rks = [1, 4, 6, 6, 1]
    for orth in (False, True):
        dr = _draws_for(dims, rks, None)
        Wt, st = O.tt_recursive_sketch_tuple([x, y], rks, dr, orth)
        We, se = O.tt_recursive_sketch(xy, rks, dr, orth)
        assert st == se
        for k in range(N + 1):
            assert Wt[k].shape == (x.rks[k], y.rks[k], se[k])
            ex = We[k].reshape(x.rks[k], y.rks[k], -1, order="F")
            np.testing.assert_allclose(Wt[k], ex, rtol=1e-11, atol=1e-12 * np.abs(ex).max())
        rt = O.ttrand_rounding_tuple([x, y], rks, dr, orth)
        re = O.ttrand_rounding(xy, rks, dr, orth)
        assert rt.rks == re.rks
        np.testing.assert_allclose(O.ttv_to_tensor(rt), O.ttv_to_tensor(re), rtol=1e-10, atol=1e-11)
        assert O.rel_err(xy, rt) < 1e-11
    # boundary sketch of the product
    bv, _ = O.sketch_structure(dims, [1] * (N + 1))
    pc = -(-7 // bv[0])
    dr = [RNG.standard_normal(bv[k + 1] * dims[k] * bv[k] * pc) for k in range(N)]
    st, sa = O.tt_sketch_tuple([x, y], 7, dr, True)
    se, sb = O.tt_sketch(xy, 7, dr, True)
    assert sa == sb
    np.testing.assert_allclose(st, se, rtol=1e-11, atol=1e-12 * np.abs(se).max())


def _fwd_draws(dims, rks, block_rks=None, p=None, oversampling=1, g=None):
    g = g or RNG
    bv, pp = O.sketch_structure_forward(dims, rks, block_rks, p, oversampling)
    return [g.standard_normal(bv[k] * dims[k] * bv[k + 1] * pp[k + 1]) for k in range(len(dims))]


@pytest.mark.parametrize("orthogonal,oversampling", [(False, 1), (True, 1), (False, 1.5)])
def test_forward_tuple_sketch_factored(orthogonal, oversampling):
    """tt_recursive_sketch(NTuple; reverse=false) (sketches.jl:539-604, 1332-1478):
    the factored forward sketch equals the forward sketch of the explicit product in
    tuple order (first factor fastest), which is pinned by the dense forward sketch."""
    dims = (2, 3, 2, 2)
    N = len(dims)
    x = rand_tt(dims, [1, 2, 3, 2, 1], 81)
    y = rand_tt(dims, [1, 2, 2, 3, 1], 82)
    z = rand_tt(dims, [1, 1, 2, 2, 1], 83)
    for fs in ([x, y], [x, y, z]):
        xy = O.hadamard(fs, first_fastest=True)
        rks = [1, 4, 6, 6, 4]
        dr = _fwd_draws(dims, rks, None, oversampling=oversampling)
        Wt, st = O.tt_recursive_sketch_tuple_forward(fs, rks, dr, orthogonal, oversampling=oversampling)
        We, se = O.tt_recursive_sketch_forward(xy, rks, dr, orthogonal, oversampling=oversampling)
        assert st == se
        for k in range(N + 1):
            assert Wt[k].shape == tuple(t.rks[k] for t in fs) + (se[k],)
            ex = We[k].reshape(tuple(t.rks[k] for t in fs) + (-1,), order="F")
            np.testing.assert_allclose(Wt[k], ex, rtol=1e-11, atol=1e-12 * np.abs(ex).max())


def test_forward_boundary_sketches_dense():
    """tt_sketch(...; reverse=false) (sketches.jl:679-694, 796-811, 894-900): p =
    cld(s, bv_{N+1}) blocks at every core, W_{N+1} / sqrt(p) = vec(X)^T Omega_{1:N} /
    sqrt(p) with the dense chain of blocks; the operator form equals the plain one
    on the explicit H x, the tuple form equals the plain one on the product."""
    dims = (2, 3, 2, 2)
    N = len(dims)
    x = rand_tt(dims, [1, 2, 4, 3, 1], 84)
    s = 7
    bv, _ = O.sketch_structure_forward(dims, [1] * (N + 1))
    pc = -(-s // bv[N])
    for orthogonal in (False, True):
        dr = [RNG.standard_normal(bv[k] * dims[k] * bv[k + 1] * pc) for k in range(N)]
        got, sa = O.tt_sketch(x, s, dr, orthogonal, reverse=False)
        assert sa == bv[N] * pc and sa >= s
        pv = [1] + [pc] * N
        Sb = [O.generate_sketch_blocks(dr[k], bv[k], dims[k], bv[k + 1], pc, orthogonal) for k in range(N)]
        ref = (_left_part(x, N).T @ _dense_sketch_fwd(Sb, dims, pv, N))[0] / math.sqrt(pc)
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    H = O.hubbard_1d_mpo(3, 1.0, 4.0)
    xh = rand_tt(H.dims, [1, 3, 4, 1], 85)
    bvh, _ = O.sketch_structure_forward(H.dims, [1] * 4)
    pch = -(-9 // bvh[3])
    dr = [RNG.standard_normal(bvh[k] * H.dims[k] * bvh[k + 1] * pch) for k in range(3)]
    got, _ = O.tt_sketch(xh, 9, dr, False, H=H, reverse=False)
    ref, _ = O.tt_sketch(O.mpo_apply(H, xh), 9, dr, False, reverse=False)
    np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-12 * np.abs(ref).max())
    y = rand_tt(dims, [1, 2, 2, 2, 1], 86)
    dr = [RNG.standard_normal(bv[k] * dims[k] * bv[k + 1] * pc) for k in range(N)]
    got, sa = O.tt_sketch_tuple([x, y], s, dr, True, reverse=False)
    ref, sb = O.tt_sketch(O.hadamard([x, y], first_fastest=True), s, dr, True, reverse=False)
    assert sa == sb
    np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-12 * np.abs(ref).max())


def test_forward_combined_sketch_parts():
    """tt_combined_sketch(...; reverse=false) (sketches.jl:930-989): the right
    boundary rank raised to max(rks_{N+1}, min(s, rks_N)); the forward recursive
    sketch's W_{N+1} alone when it has s entries, else [sqrt(l/s) W_{N+1};
    sqrt(le/s) tt_sketch(seed + 1)]; W_{N+1} replaced by ones."""
    dims = (2, 3, 2, 2)
    N = len(dims)
    x = rand_tt(dims, [1, 2, 4, 3, 1], 87)
    rks = [1, 3, 4, 3, 1]
    rk1 = rks[:N] + [max(1, min(2, rks[N - 1]))]
    dr = _fwd_draws(dims, rk1)
    cs, W, srk = O.tt_combined_sketch(x, rks, 2, dr, None, False, reverse=False)
    Wr, sr = O.tt_recursive_sketch_forward(x, rk1, dr, False)
    np.testing.assert_array_equal(cs, O.jreshape(Wr[N], -1))
    assert srk == sr[:N] + [1] and W[N].shape == (1, 1)
    assert all(np.array_equal(a, b) for a, b in zip(W[:N], Wr[:N]))
    s = 13
    rk2 = rks[:N] + [max(1, min(s, rks[N - 1]))]
    dr = _fwd_draws(dims, rk2)
    Wr, sr = O.tt_recursive_sketch_forward(x, rk2, dr, False)
    bvs, _ = O.sketch_structure_forward(dims, [1] * (N + 1))
    pe = -(-(s - sr[N]) // bvs[N])
    de = [RNG.standard_normal(bvs[k] * dims[k] * bvs[k + 1] * pe) for k in range(N)]
    ex, le = O.tt_sketch(x, s - sr[N], de, False, reverse=False)
    cs, W, srk = O.tt_combined_sketch(x, rks, s, dr, de, False, reverse=False)
    tot = sr[N] + le
    assert len(cs) == tot and tot >= s
    np.testing.assert_allclose(cs[: sr[N]], math.sqrt(sr[N] / tot) * O.jreshape(Wr[N], -1), rtol=1e-15)
    np.testing.assert_allclose(cs[sr[N]:], math.sqrt(le / tot) * ex, rtol=1e-15)


# ---------------------------------------------------------------- RNG

def test_philox_known_answers():
    def kat(c, k):
        r = P.philox4x32_10(*[np.uint64(v) for v in c], np.uint64(k[0]), np.uint64(k[1]))
        return [int(v) for v in r]
    assert kat([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert kat([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert kat([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_normal_draws_statistics():
    g = P.normal_draws(1234, 0, 400000)
    assert abs(g.mean()) < 0.01 and abs(g.var() - 1.0) < 0.01
    assert abs(np.mean(g ** 4) - 3.0) < 0.05
    assert not np.array_equal(g[:100], P.normal_draws(1234, 1, 100))
    np.testing.assert_array_equal(g[50:60], P.normal_draws(1234, 0, 10, start=50))
