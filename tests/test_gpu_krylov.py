"""Sketched Krylov basis and sketched Rayleigh-Ritz on the device
(src/tt_solvers.jl:354-689): the host mirror's expand_basis_, sketched_rr and
sketched_rayleigh_ritz, whose sketches, randomized and SVD roundings and inner
products run through the C-ABI, against the oracle fed the device's own sketch
blocks.

Tolerances: FP64 paths, 1e-12 relative for the sketches and the new basis
vectors (represented tensor, Frobenius norm; measured ~5e-15), 1e-10 for the
stored partial sketches W_k (compared through the gauge-free product with the
vector's left part) and 1e-9 for Ritz vectors, which pass through a truncation at
tol = res_sketch^2/2 placed by two different SVDs.  The library's sketch blocks do not depend
on the block count (block j of core k is the same draw whatever p), so one
provider per seed with a generous p serves every sketch structure.
"""
import sys

import numpy as np
import pytest

import oracle as O
from helpers import device_blocks_provider, to_api, to_oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def T():
    import ttround
    ttround.default_handle()
    return ttround


def api_op(H):
    import ttround as T
    return T.TToperator(H.N, [c.copy() for c in H.cores], H.dims, list(H.rks))


def provider(dims, seed, orthogonal, block_rks, P=96):
    N = len(dims)
    return device_blocks_provider(dims, [1] * (N + 1), orthogonal=orthogonal, seed=seed, block_rks=block_rks,
                                  p=[P] * N + [1])


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def left_part(x, k):
    """Cores 0..k-1 contracted: the (n_0 .. n_{k-1}) x r_k matrix."""
    M = np.ones((1, 1))
    for j in range(k):
        c = x.cores[j]  # (n, r, r')
        M = np.einsum("pa,iab->pib", M, c).reshape(-1, c.shape[2], order="F")
    return M


def start_windows_oracle(b0, rmax, s, dr, de, orthogonal, block_rks):
    N = b0.N
    rks = [2 * rmax] * (N + 1)
    rks[N] = 1
    sk, W, _ = O.tt_combined_sketch(b0, rks, s, dr, de, orthogonal, block_rks)
    beta = np.linalg.norm(sk)
    for i in range(1, b0.ot.index(0) + 1):
        W[i] = W[i] / beta
    return [O.tt_scale(1 / beta, b0)], [W], [sk / beta]


def start_windows_device(T, b0, rmax, s, orthogonal, block_rks, seed):
    sk, W, _ = T.tt_combined_sketch(to_api(b0), 2 * rmax, s, orthogonal=orthogonal, seed=seed, block_rks=block_rks)
    beta = np.linalg.norm(sk)
    for i in range(1, b0.ot.index(0) + 1):
        W[i] = W[i] / beta
    return [to_api(b0) / beta], [W], [sk / beta]


@pytest.mark.parametrize("orthogonal", [False, True])
def test_expand_basis_vs_oracle(T, orthogonal):
    L, rmax, s, seed, bk = 5, 6, 20, 31, 4
    say(f"case: expand_basis_ L={L} rmax={rmax} s={s} orthogonal={orthogonal}, 3 steps, k_trunc=2")
    H = O.hubbard_1d_mpo(L, 1.0, 4.0)
    Ha = api_op(H)
    b0 = O.rand_tt(H.dims, [1, 3, 5, 6, 4, 1], np.random.default_rng(140))
    dr, de = provider(H.dims, seed, orthogonal, bk), provider(H.dims, seed + 1, orthogonal, bk)
    Bo, Wo, So = start_windows_oracle(b0, rmax, s, dr, de, orthogonal, bk)
    Bd, Wd, Sd = start_windows_device(T, b0, rmax, s, orthogonal, bk, seed)
    assert rel(Sd[0], So[0]) < 1e-12
    for step in range(3):
        ro = O.expand_basis(H, Bo, Wo, So, rmax, s, dr, de, orthogonal, bk, k_trunc=2)
        rd = T.expand_basis_(Ha, Bd, Wd, Sd, rmax, s, orthogonal=orthogonal, block_rks=bk, seed=seed, k_trunc=2)
        e_hb = rel(rd[2], ro[2])
        e_b = O.rel_err(ro[0], to_oracle(rd[0]))
        e_s = rel(rd[1], ro[1])
        say(f"  step {step}: ranks {rd[0].ttv_rks}  sketch(Hb) {e_hb:.2e}  b_new {e_b:.2e}  sketch(b_new) {e_s:.2e}")
        assert rd[0].ttv_rks == ro[0].rks
        assert e_hb < 1e-12 and e_b < 1e-12 and e_s < 1e-12
        assert len(Bd) == len(Wd) == len(Sd) == len(Bo) == min(2, step + 2)
        # the stored sketches W_k in the gauge of each b_new: the left part times W_k
        # is the gauge-free (n_1..n_{k-1}) x s_k sketch of the partial unfolding
        xd, xo = to_oracle(Bd[-1]), Bo[-1]
        for k in range(1, L):
            assert rel(left_part(xd, k) @ Wd[-1][k], left_part(xo, k) @ Wo[-1][k]) < 1e-10
        # sketch-orthogonality of the new vector to the window it was built from
        assert abs(np.linalg.norm(rd[1]) - 1) < 1e-13


@pytest.mark.parametrize("stable", [False, True])
def test_sketched_rr_invariant_subspace(T, stable):
    """Exact eigenvectors of H as the basis: the Ritz values are H's eigenvalues
    whatever the sketch, the Ritz vector the lowest eigenvector."""
    L, rmax, seed, bk = 3, 8, 5, 4
    say(f"case: sketched_rr on eigenvectors, stable={stable}")
    H = O.hubbard_1d_mpo(L, 1.0, 4.0)
    w, V = np.linalg.eigh(O.hubbard_1d_dense(L, 1.0, 4.0))
    from test_oracle_krylov import dense_to_tt
    pick = [5, 0, 9]
    B = [dense_to_tt(np.reshape(V[:, i], H.dims, order="F"), H.dims) for i in pick]
    Ba = [to_api(b) for b in B]
    Ha = api_op(H)
    C = np.column_stack([T.tt_combined_sketch(b, 2 * rmax, 16, seed=seed, block_rks=bk)[0] for b in Ba])
    D = np.column_stack([T.tt_combined_sketch(Ha, b, 2 * rmax, 16, seed=seed, block_rks=bk)[0] for b in Ba])
    lam, X, rs, rsamp = T.sketched_rr(Ha, Ba, C, D, rmax, block_rks=bk, seed=seed, stable=stable, seed_samp=77,
                                      verbose=False)
    say(f"  lambdas {lam}  res_sketch {rs:.2e}  res_sample {rsamp:.2e}")
    np.testing.assert_allclose(lam, np.sort(w[pick]), atol=1e-10)
    x = O.ttv_to_tensor(to_oracle(X)).reshape(-1, order="F")
    assert abs(abs(x @ V[:, 0]) - 1) < 1e-10
    assert rs < 1e-10 and rsamp < 1e-9
    # the oracle on the same draws and sketches
    N = L
    lo, Xo, rso, rsao = O.sketched_rr(H, B, C, D, rmax, provider(H.dims, seed, True, bk),
                                      provider(H.dims, 77, True, bk), True, bk, stable=stable)
    np.testing.assert_allclose(lam, lo, rtol=1e-12, atol=1e-12)
    assert O.rel_err(Xo, to_oracle(X)) < 1e-9


def test_sketched_rr_krylov_basis_vs_oracle(T):
    """A sketch-orthogonal Krylov basis (built by the oracle on the device's draws):
    Ritz values, Ritz vector and residuals against the oracle."""
    L, rmax, s, seed, bk, d = 4, 8, 24, 12, 4, 6
    say(f"case: sketched_rr on a Krylov basis of {d} vectors, L={L}")
    H = O.hubbard_1d_mpo(L, 1.0, 4.0)
    Ha = api_op(H)
    b0 = O.rand_tt(H.dims, [1, 4, 6, 4, 1], np.random.default_rng(9))
    dr, de = provider(H.dims, seed, True, bk), provider(H.dims, seed + 1, True, bk)
    Bw, Ww, Sw = start_windows_oracle(b0, rmax, s, dr, de, True, bk)
    B, Cs, Ds = [Bw[0]], [Sw[0]], []
    for j in range(d):
        b_new, sk_new, sk_hb = O.expand_basis(H, Bw, Ww, Sw, rmax, s, dr, de, True, bk, k_trunc=d)
        B.append(b_new)
        Cs.append(sk_new)
        Ds.append(sk_hb)
    B, C, D = B[:d], np.column_stack(Cs[:d]), np.column_stack(Ds[:d])
    lo, Xo, rso, rsao = O.sketched_rr(H, B, C, D, rmax, provider(H.dims, seed, True, bk),
                                      provider(H.dims, 4242, True, bk), True, bk)
    lam, X, rs, rsa = T.sketched_rr(Ha, [to_api(b) for b in B], C, D, rmax, block_rks=bk, seed=seed,
                                    seed_samp=4242, verbose=False)
    say(f"  lambda_1 {lam[0]:.10f} (oracle {lo[0]:.10f})  res_sketch {rs:.3e}  res_sample {rsa:.3e} (oracle {rsao:.3e})")
    np.testing.assert_allclose(lam, lo, rtol=1e-12, atol=1e-12)
    assert abs(rs - rso) <= 1e-10 * max(1.0, rso)
    if np.isfinite(rsao):
        assert O.rel_err(Xo, to_oracle(X)) < 1e-9
        assert abs(rsa - rsao) <= 1e-8 * rsao
    else:  # res_sketch^2/2 above the Ritz vector's norm: both zeroed, both NaN
        assert not np.isfinite(rsa)


def test_sketched_rayleigh_ritz_device(T):
    """The whole driver on the device: every finite Rayleigh quotient above the
    dense ground-state energy, and a Krylov dimension of 20 on the 64-dimensional
    3-site chain reaches it; the first steps agree with the oracle."""
    L, d, rmax, seed, bk = 3, 20, 4, 3, 4
    say(f"case: sketched_rayleigh_ritz L={L} d={d} rmax={rmax}")
    H = O.hubbard_1d_mpo(L, 1.0, 4.0)
    E0 = np.linalg.eigvalsh(O.hubbard_1d_dense(L, 1.0, 4.0))[0]
    b0 = O.rand_tt(H.dims, [1, 4, 4, 1], np.random.default_rng(21))
    samp = list(range(40, 40 + d))
    e, psi, hist = T.sketched_rayleigh_ritz(api_op(H), to_api(b0), d, rmax, orthogonal=False, block_rks=bk,
                                            seed=seed, sample_seeds=samp, verbose=False)
    finite = [h for h in hist if np.isfinite(h["e_true"])]
    say(f"  E0 {E0:.12f}  e_sketch {e:.12f}  e_true {hist[-1]['e_true']:.12f}  NaN steps {d - len(finite)}")
    assert len(hist) == d and [h["iter"] for h in hist] == list(range(1, d + 1))
    assert len(finite) >= d - 4 and finite[-1] is hist[-1]
    for h in finite:
        assert h["e_true"] >= E0 - 1e-9
    assert abs(hist[-1]["e_true"] - E0) < 1e-6 * abs(E0)
    assert abs(e - E0) < 1e-4 * abs(E0)
    assert abs(T.norm(psi) - 1) < 1e-12
    dims = H.dims
    draws_for = lambda sd: provider(dims, sd, False, bk)  # noqa: E731
    eo, psio, histo = O.sketched_rayleigh_ritz(H, b0, d, rmax, draws_for, orthogonal=False, block_rks=bk, seed=seed,
                                               sample_seeds=samp)
    for hd, ho in zip(hist[:4], histo[:4]):
        for key in ("e_sketch", "e_true", "res_sketch"):
            a, b = hd[key], ho[key]
            assert (not np.isfinite(a) and not np.isfinite(b)) or abs(a - b) <= 1e-8 * max(1.0, abs(b)), (key, hd, ho)
