"""Pins of the oracle's sketched Krylov functions (CPU only).

  expand_basis          -> with sketch ranks above the true ranks the randomized
                           rounding and the rmax rounding are exact, so the new
                           vector is the dense (H b_prev - sum_i h_i b_i) / beta,
                           beta = the norm of its boundary sketch; its boundary
                           sketch is orthogonal to the window's sketches (the
                           least-squares normal equations of h); the sliding window
                           keeps k_trunc entries (src/tt_solvers.jl:354-455)
  sketched_rr           -> on a basis of exact eigenvectors of H, C^+ D is the
                           diagonal of their eigenvalues whatever the sketch: the
                           Ritz values are H's eigenvalues in ascending order, the
                           Ritz vector is the lowest eigenvector, both residuals
                           vanish; both branches (stable and column-pivoted)
                           (src/tt_solvers.jl:466-514)
  sketched_rayleigh_ritz -> every true Rayleigh quotient lies above the dense
                           ground-state energy, and a Krylov dimension covering
                           the space reaches it (src/tt_solvers.jl:548-620)
A slip (a sign of h, a dropped term of the sweep, an unsorted spectrum, D and C
swapped, a missing normalisation) changes the dense quantities these compare.
"""
import numpy as np
import pytest

import oracle as O


def rand_tt(dims, rks, seed):
    return O.rand_tt(dims, rks, np.random.default_rng(seed))


def draws_for(seed, N=8, n=1 << 15):
    rng = np.random.default_rng(1000 + seed)
    return [rng.standard_normal(n) for _ in range(N)]


def dense_to_tt(T, dims):
    """Exact TT of a dense tensor by sequential SVDs (site 1 fastest)."""
    d = len(dims)
    cores, r = [], 1
    M = np.reshape(T, (r * dims[0], -1), order="F")
    for k in range(d - 1):
        M = np.reshape(M, (r * dims[k], -1), order="F")
        u, s, vt = np.linalg.svd(M, full_matrices=False)
        rk = int(np.sum(s > 1e-13 * s[0]))
        core = np.reshape(u[:, :rk], (r, dims[k], rk), order="F").transpose(1, 0, 2)
        cores.append(np.ascontiguousarray(core))
        M = s[:rk, None] * vt[:rk]
        r = rk
    cores.append(np.reshape(M, (r, dims[-1], 1), order="F").transpose(1, 0, 2).copy())
    rks = [1] + [c.shape[2] for c in cores]
    return O.TT(cores, tuple(dims), rks, [0] * d)


def dense_H(L):
    return O.hubbard_1d_dense(L, 1.0, 4.0)


def test_expand_basis_exact_step():
    L, rmax, s = 4, 16, 12
    H = O.hubbard_1d_mpo(L, 1.0, 4.0)
    Hd = dense_H(L)
    dims = H.dims
    b0 = rand_tt(dims, [1, 2, 3, 2, 1], 7)
    rks = [2 * rmax] * (L + 1)
    rks[L] = 1
    dr, de = draws_for(5), draws_for(6)
    sk1, W1, _ = O.tt_combined_sketch(b0, rks, s, dr, de, False, 4)
    beta = np.linalg.norm(sk1)
    root = b0.ot.index(0)
    for i in range(1, root + 1):
        W1[i] = W1[i] / beta
    Bw, Ww, Sw = [O.tt_scale(1 / beta, b0)], [W1], [sk1 / beta]
    for step in range(3):
        prev = [O.ttv_to_tensor(b).reshape(-1, order="F") for b in Bw]
        Swin = np.column_stack(Sw)
        b_new, sk_new, sk_Hb = O.expand_basis(H, Bw, Ww, Sw, rmax, s, dr, de, False, 4, k_trunc=2)
        # h from the same least-squares problem, then the dense step
        h = np.linalg.lstsq(Swin, sk_Hb, rcond=None)[0]
        v = Hd @ prev[-1] - np.column_stack(prev) @ h
        got = O.ttv_to_tensor(b_new).reshape(-1, order="F")
        scale = np.linalg.norm(v) / np.linalg.norm(got)
        np.testing.assert_allclose(got * scale, v, rtol=0, atol=1e-11 * np.linalg.norm(v))
        # the boundary sketch of b_new: unit norm, orthogonal to the window's sketches
        assert abs(np.linalg.norm(sk_new) - 1) < 1e-13
        assert np.max(np.abs(Swin.T @ sk_new)) < 1e-11
        # beta is the norm of the sketch of the unnormalised vector: sketch is linear
        sk_v = (sk_Hb - Swin @ h)
        np.testing.assert_allclose(sk_new * np.linalg.norm(sk_v), sk_v, atol=1e-11 * np.linalg.norm(sk_v))
        assert len(Bw) == len(Ww) == len(Sw) == min(2, step + 2)
        assert Bw[-1] is b_new


@pytest.mark.parametrize("stable", [False, True])
def test_sketched_rr_invariant_subspace(stable):
    L, rmax = 3, 8
    H = O.hubbard_1d_mpo(L, 1.0, 4.0)
    Hd = dense_H(L)
    w, V = np.linalg.eigh(Hd)
    dims = H.dims
    pick = [5, 0, 9]  # eigenvectors, deliberately not in ascending order
    B = [dense_to_tt(np.reshape(V[:, i], dims, order="F"), dims) for i in pick]
    rks = [2 * rmax] * (L + 1)
    rks[L] = 1
    dr, de = draws_for(11), draws_for(12)
    s = 16
    C = np.column_stack([O.tt_combined_sketch(b, rks, s, dr, de, False, 4)[0] for b in B])
    D = np.column_stack([O.tt_combined_sketch(b, rks, s, dr, de, False, 4, H=H)[0] for b in B])
    lam, X, rs, rsamp = O.sketched_rr(H, B, C, D, rmax, draws_for(13), draws_for(14), False, 4, stable=stable)
    np.testing.assert_allclose(lam, np.sort(w[pick]), atol=1e-10)
    x = O.ttv_to_tensor(X).reshape(-1, order="F")
    assert abs(abs(x @ V[:, 0]) - 1) < 1e-10
    assert rs < 1e-10 and rsamp < 1e-9


def test_sketched_rayleigh_ritz_variational():
    L, d = 3, 20
    H = O.hubbard_1d_mpo(L, 1.0, 4.0)
    E0 = np.linalg.eigvalsh(dense_H(L))[0]
    b0 = rand_tt(H.dims, [1, 4, 4, 1], 21)
    e, psi, hist = O.sketched_rayleigh_ritz(H, b0, d, 4, draws_for, orthogonal=False, block_rks=4, seed=3,
                                            sample_seeds=list(range(40, 40 + d)))
    assert len(hist) == d and [h["iter"] for h in hist] == list(range(1, d + 1))
    # an early Ritz vector is NaN when res_sketch^2/2 exceeds its norm (the
    # rounding tolerance of src/tt_solvers.jl:503 zeroes it); every other one is
    # a normalised vector, so its Rayleigh quotient is at least E0
    finite = [h for h in hist if np.isfinite(h["e_true"])]
    assert len(finite) >= d - 4 and finite[-1] is hist[-1]
    for h in finite:
        assert h["e_true"] >= E0 - 1e-9
    assert abs(hist[-1]["e_true"] - E0) < 1e-6 * abs(E0), (hist[-1], E0)
    assert abs(e - E0) < 1e-4 * abs(E0)
    assert abs(O.tt_norm(psi) - 1) < 1e-12
