"""Pins of the oracle's sweep directions, operator rounding and truncating
randomized rounding (CPU only).

  tt_rounding(direction="right") -> dense TT-SVD of the unfoldings from left to
                                    right with the same floor rule and rmax
                                    (src/tt_rounding.jl:215-264)
  tto_to_ttv / ttv_to_tto        -> round trip, and the dense vectorised operator
                                    equals the dense operator with (i_k, j_k)
                                    merged, i_k fastest (src/tt_tools.jl:603-626)
  tto_rounding                   -> dense TT-SVD of that vectorisation; H + H of the
                                    Hubbard MPO rounds to 2H with H's ranks
                                    (src/tt_rounding.jl:275-277)
  ttrand_rounding*, truncating   -> dense brute force of the recursive projection:
                                    Q_k from the thin QR of
                                    (Q_{<k}^T x I) X_(k) Omega_{k+1:N}, result
                                    Q_{<N} Q_{<N}^T X (src/tt_randtools.jl:112-154,
                                    223-276, 285-354), with sketch ranks far below the
                                    true ranks, so nothing is recovered exactly
A plausible slip (a transposed unfolding, the wrong sweep order, a missing
projection, a swapped (i, j) merge) changes the dense results these compare.
"""
import math

import numpy as np
import pytest

import oracle as O
from oracle import philox as P


def rand_tt(dims, rks, seed):
    return O.rand_tt(dims, rks, np.random.default_rng(seed))


def dense_tt_svd(T, dims, tol, rmax, direction):
    """Sequential truncation of the dense unfoldings with the floor rule:
    bonds d-1 .. 1 (direction "left") or 1 .. d-1 ("right")."""
    T = T.copy()
    d = len(dims)
    ranks = [1] * (d + 1)
    bonds = range(d - 1, 0, -1) if direction == "left" else range(1, d)
    for b in bonds:  # bond between modes b | b+1 (1-based), i.e. rows = n_1..n_b
        rows = int(np.prod(dims[:b]))
        M = np.reshape(T, (rows, -1), order="F")
        u, s, vt = np.linalg.svd(M, full_matrices=False)
        _, k = O.floor_rule(s, tol)
        k = min(k, rmax, int(np.sum(s > 0)))
        T = np.reshape((u[:, :k] * s[:k]) @ vt[:k], T.shape, order="F")
        ranks[b] = k
    return T, ranks


@pytest.mark.parametrize("tol,rmax", [(1e-12, 2 ** 14), (1e-2, 2 ** 14), (1e-12, 3), (0.3, 5)])
def test_tt_rounding_right_matches_dense_svd(tol, rmax):
    dims = (2, 3, 2, 3, 2)
    x = O.tt_add(rand_tt(dims, 3, 5), O.tt_scale(0.05, rand_tt(dims, 4, 6)))
    X = O.ttv_to_tensor(x)
    Td, ranks = dense_tt_svd(X, dims, tol, rmax, "right")
    y = O.tt_rounding(x, tol=tol, rmax=rmax, direction="right")
    assert y.rks == ranks
    np.testing.assert_allclose(O.ttv_to_tensor(y), Td, rtol=1e-11, atol=1e-11 * np.abs(X).max())
    assert y.ot == [1] * (len(dims) - 1) + [0]
    # left-orthogonal cores
    for k in range(len(dims) - 1):
        c = y.cores[k]
        np.testing.assert_allclose(sum(c[i].T @ c[i] for i in range(c.shape[0])), np.eye(c.shape[2]), atol=1e-12)


def test_directions_differ_when_truncating():
    """The two sweep orders truncate different tensors; a restatement that ran
    the same sweep for both would pass the dense test of only one of them."""
    dims = (2, 3, 2, 3, 2)
    x = O.tt_add(rand_tt(dims, 3, 5), O.tt_scale(0.05, rand_tt(dims, 4, 6)))
    X = O.ttv_to_tensor(x)
    Tl, _ = dense_tt_svd(X, dims, 0.3, 5, "left")
    Tr, _ = dense_tt_svd(X, dims, 0.3, 5, "right")
    assert np.linalg.norm(Tl - Tr) > 1e-6 * np.linalg.norm(X)


def _rand_tto(dims, rks, seed):
    g = np.random.default_rng(seed)
    return O.TTO([g.standard_normal((n, n, rks[k], rks[k + 1])) for k, n in enumerate(dims)], dims, list(rks))


def _vectorised_dense(A):
    """tto_to_tensor(A) (rows i, columns j, column-major multi-indices) as the
    dense tensor of tto_to_ttv(A): mode k is (i_k, j_k) with i_k fastest."""
    d, n = A.N, A.dims
    M = O.tto_to_tensor(A)
    T = np.reshape(M, tuple(n) + tuple(n), order="F")  # axes i_1..i_d, j_1..j_d
    axes = [a for k in range(d) for a in (k, d + k)]   # i_1, j_1, i_2, j_2, ...
    T = np.transpose(T, axes)
    return np.reshape(T, tuple(m * m for m in n), order="F")


def test_tto_ttv_roundtrip_and_layout():
    dims = (2, 3, 2)
    A = _rand_tto(dims, [1, 2, 3, 1], 70)
    v = O.tto_to_ttv(A)
    assert v.dims == (4, 9, 4) and v.rks == A.rks
    np.testing.assert_allclose(O.ttv_to_tensor(v), _vectorised_dense(A), rtol=1e-13, atol=1e-13)
    B = O.ttv_to_tto(v)
    assert B.dims == A.dims and B.rks == A.rks
    for a, b in zip(A.cores, B.cores):
        np.testing.assert_array_equal(a, b)
    # a vectorisation that merged (j, i) instead would differ for non-symmetric cores
    c = A.cores[1]
    assert np.linalg.norm(c - c.transpose(1, 0, 2, 3)) > 0.1 * np.linalg.norm(c)
    with pytest.raises(AssertionError):
        O.ttv_to_tto(rand_tt((3, 4), 2, 71))


@pytest.mark.parametrize("tol,rmax", [(1e-12, 2 ** 14), (1e-12, 2), (0.2, 2 ** 14)])
def test_tto_rounding_matches_dense(tol, rmax):
    dims = (2, 3, 2, 2)
    A = O.TTO([a + 0.1 * b for a, b in zip(_rand_tto(dims, [1, 3, 4, 3, 1], 72).cores,
                                          _rand_tto(dims, [1, 3, 4, 3, 1], 73).cores)], dims, [1, 3, 4, 3, 1])
    Vd = _vectorised_dense(A)
    Td, ranks = dense_tt_svd(Vd, tuple(n * n for n in dims), tol, rmax, "left")
    R = O.tto_rounding(A, tol=tol, rmax=rmax)
    assert R.rks == ranks
    np.testing.assert_allclose(_vectorised_dense(R), Td, rtol=1e-11, atol=1e-11 * np.abs(Vd).max())
    assert R.dims == dims


def test_tto_rounding_hubbard_sum():
    """H + H (MPO ranks doubled to 12) rounds to 2H with H's ranks (6 inside)."""
    L = 4
    H = O.hubbard_1d_mpo(L, 1.0, 4.0)
    HH = O.TTO([np.zeros((4, 4, 2 * H.rks[k] if 0 < k else 1, 2 * H.rks[k + 1] if k + 1 < L else 1))
                for k in range(L)], H.dims, [1] + [2 * r for r in H.rks[1:-1]] + [1])
    for k in range(L):  # block-diagonal direct sum of H with itself, as tt_add does for vectors
        c, R0, R1 = H.cores[k], H.rks[k], H.rks[k + 1]
        if k == 0:
            HH.cores[k][:, :, 0, :R1] = c[:, :, 0, :]
            HH.cores[k][:, :, 0, R1:] = c[:, :, 0, :]
        elif k == L - 1:
            HH.cores[k][:, :, :R0, 0] = c[:, :, :, 0]
            HH.cores[k][:, :, R0:, 0] = c[:, :, :, 0]
        else:
            HH.cores[k][:, :, :R0, :R1] = c
            HH.cores[k][:, :, R0:, R1:] = c
    np.testing.assert_allclose(O.tto_to_tensor(HH), 2 * O.hubbard_1d_dense(L, 1.0, 4.0), atol=1e-12)
    R = O.tto_rounding(HH, tol=1e-12)
    assert max(R.rks) <= 6 and R.rks[1:-1] == [min(6, r) for r in O.tto_rounding(H, tol=1e-12).rks[1:-1]]
    np.testing.assert_allclose(O.tto_to_tensor(R), 2 * O.hubbard_1d_dense(L, 1.0, 4.0), atol=1e-11)


# ---------------------------------------------------------------- truncating randomized rounding

def _draws_for(dims, rks, block_rks, seed=1234):
    bv, p = O.sketch_structure(dims, rks, block_rks)
    return [P.normal_draws(seed, k, bv[k + 1] * dims[k] * bv[k] * p[k]) for k in range(len(dims))]


def _omega_from_blocks(dims, bv, p, Sb, k):
    """Omega_{k:N}[(i_k..i_N), (b, j)] = sum over the chain of blocks of block j,
    b the row index of block j at core k (column-major over (b, j))."""
    N = len(dims)
    cols = []
    for j in range(p[k]):
        om = None
        for q in range(N - 1, k - 1, -1):
            Sq = Sb[q][:, :, :, j]  # (bv_{q+1}, n_q, bv_q): chain of block j
            if q == N - 1:
                om = Sq[0]  # (n, bv)
            else:
                t = np.einsum("bic,rb->irc", Sq, om)  # (n_q, rest, bv_q)
                om = t.reshape(-1, Sq.shape[2], order="F")
        cols.append(om)
    return np.concatenate(cols, axis=1)


def brute_force_projection(X, dims, Om, weights=None):
    """Dense restatement of randomize-then-orthogonalize: Q_k = thin QR of
    (Q_{<k}^T x I) X_(k) Omega_{k+1:N}; returns Q_{<N} Q_{<N}^T X and the ranks."""
    N = len(dims)
    Qlt = np.ones((1, 1))  # (n_1..n_{k-1}) x r_{k-1}
    ranks = [1] * (N + 1)
    for k in range(N - 1):
        rows = int(np.prod(dims[:k + 1]))
        Xk = np.reshape(X, (rows, -1), order="F")  # (n_1..n_k) x (n_{k+1}..n_N)
        XO = Xk @ Om[k + 1]  # (n_1..n_k) x s_{k+1}
        # (Q_{<k}^T x I_{n_k}) on rows (i_<k, i_k), i_<k fastest -> rows (i_k, a), i_k fastest
        prev = int(np.prod(dims[:k]))
        T3 = np.reshape(XO, (prev, dims[k], XO.shape[1]), order="F")
        Z = np.einsum("pa,pis->ias", Qlt, T3).reshape(dims[k] * Qlt.shape[1], -1, order="F")
        Q, _ = np.linalg.qr(Z)
        ranks[k + 1] = Q.shape[1]
        Qc = np.reshape(Q, (dims[k], Qlt.shape[1], Q.shape[1]), order="F")  # core (i_k, a, b)
        Qlt = np.einsum("pa,iab->pib", Qlt, Qc).reshape(rows, Q.shape[1], order="F")
    rows = int(np.prod(dims[:N - 1]))
    Xl = np.reshape(X, (rows, -1), order="F")
    return np.reshape(Qlt @ (Qlt.T @ Xl), X.shape, order="F"), ranks


def _omegas(dims, rks, draws, orthogonal, block_rks):
    bv, p = O.sketch_structure(dims, rks, block_rks)
    Sb = [O.generate_sketch_blocks(draws[k], bv[k + 1], dims[k], bv[k], p[k], orthogonal) for k in range(len(dims))]
    return [None] + [_omega_from_blocks(dims, bv, p, Sb, k) for k in range(1, len(dims))]


@pytest.mark.parametrize("orthogonal,block_rks", [(False, None), (True, None), (False, 2)])
def test_ttrand_rounding_truncating_matches_projection(orthogonal, block_rks):
    dims = (3, 2, 3, 2, 3)
    x = O.tt_add(rand_tt(dims, 5, 80), O.tt_scale(0.3, rand_tt(dims, 4, 81)))
    rks = [1, 2, 3, 3, 2, 1]  # far below the true ranks (up to 9)
    draws = _draws_for(dims, rks, block_rks, seed=55)
    y = O.ttrand_rounding(x, rks, draws, orthogonal, block_rks)
    X = O.ttv_to_tensor(x)
    ref, ranks = brute_force_projection(X, dims, _omegas(dims, rks, draws, orthogonal, block_rks))
    assert y.rks == ranks
    Y = O.ttv_to_tensor(y)
    # truncating for real: the rounding error is large, and the projection pins it
    assert np.linalg.norm(Y - X) > 0.05 * np.linalg.norm(X)
    np.testing.assert_allclose(Y, ref, rtol=1e-11, atol=1e-11 * np.abs(X).max())


def test_ttrand_rounding_lincomb_truncating_matches_projection():
    dims = (2, 3, 2, 3, 2)
    ys = [rand_tt(dims, r, 90 + r) for r in (3, 4, 2)]
    al = [1.0, -0.5, 2.0]
    rks = [1, 2, 3, 3, 2, 1]
    draws = _draws_for(dims, rks, None, seed=56)
    y = O.ttrand_rounding_lincomb(al, ys, rks, draws, False)
    X = sum(a * O.ttv_to_tensor(t) for a, t in zip(al, ys))
    ref, ranks = brute_force_projection(X, dims, _omegas(dims, rks, draws, False, None))
    assert y.rks == ranks
    Y = O.ttv_to_tensor(y)
    assert np.linalg.norm(Y - X) > 0.05 * np.linalg.norm(X)
    np.testing.assert_allclose(Y, ref, rtol=1e-11, atol=1e-11 * np.abs(X).max())


def test_ttrand_rounding_op_truncating_matches_projection():
    """The implicit operator form (alpha_1 H y_1 + alpha_2 y_2) against the
    projection of the explicit dense sum."""
    L = 4
    H = O.hubbard_1d_mpo(L, 1.0, 4.0)
    dims = H.dims
    ys = [rand_tt(dims, [1, 3, 5, 3, 1], 95), rand_tt(dims, [1, 2, 4, 2, 1], 96)]
    al = [1.0, -0.7]
    rks = [1, 3, 5, 3, 1]
    draws = _draws_for(dims, rks, None, seed=57)
    y = O.ttrand_rounding_op(al, H, ys, rks, draws, False)
    X = al[0] * O.ttv_to_tensor(O.mpo_apply(H, ys[0])) + al[1] * O.ttv_to_tensor(ys[1])
    ref, ranks = brute_force_projection(X, dims, _omegas(dims, rks, draws, False, None))
    assert y.rks == ranks
    Y = O.ttv_to_tensor(y)
    assert np.linalg.norm(Y - X) > 0.01 * np.linalg.norm(X)
    np.testing.assert_allclose(Y, ref, rtol=1e-10, atol=1e-10 * np.abs(X).max())


# ---------------------------------------------------------------- forward operator sketch

@pytest.mark.parametrize("shape", [(3, 2, 4, 2, 3, 2, 5), (2, 6, 2, 4, 1, 6, 2)])
def test_contract_forwards_op_both_orders(shape):
    """Both cost-chosen orders of the forward operator contraction equal the
    direct sum W[a,b,c] = sum A[ze,al,a] H[z,ze,be,b] S[ga,z,c] V[al,be,ga]."""
    n, al, a, be, b, ga, c = shape
    g = np.random.default_rng(sum(shape))
    A, H = g.standard_normal((n, al, a)), g.standard_normal((n, n, be, b))
    S, V = g.standard_normal((ga, n, c)), g.standard_normal((al, be, ga))
    ref = np.einsum("xpa,zxqb,gzc,pqg->abc", A, H, S, V)
    np.testing.assert_allclose(O.contract_sketch_core_forwards_op(A, H, S, V), ref, rtol=1e-12, atol=1e-12)


def test_forward_op_sketch_is_the_sketch_of_the_product():
    """tt_recursive_sketch(H, x; reverse=false) equals the forward sketch of the
    explicit H x (src/tt_operations.jl:170-194, operator rank fastest in its bond)
    with the bond index reordered to x rank fastest; the operator form applies no
    oversampling (src/sketches.jl:399-402)."""
    g = np.random.default_rng(1)
    dims = (2, 3, 2, 3)
    x = rand_tt(dims, [1, 2, 3, 2, 1], 7)
    H = O.TTO([g.standard_normal((n, n, r0, r1)) for n, r0, r1 in zip(dims, [1, 2, 2, 3], [2, 2, 3, 1])], dims,
              [1, 2, 2, 3, 1])
    rks = [1, 3, 4, 3, 1]
    bv, p = O.sketch_structure_forward(dims, rks, 2)
    draws = [g.standard_normal(bv[k] * dims[k] * bv[k + 1] * p[k + 1]) for k in range(4)]
    W, s = O.tt_recursive_sketch_forward_op(H, x, rks, draws, False, 2, oversampling=2)
    We, se = O.tt_recursive_sketch_forward(O.mpo_apply(H, x), rks, draws, False, 2)
    assert s == se
    for k in range(5):
        r, R = x.rks[k], H.rks[k]
        np.testing.assert_allclose(W[k], We[k].reshape(R, r, -1, order="F").transpose(1, 0, 2), rtol=1e-12,
                                   atol=1e-12 * np.abs(We[k]).max())
