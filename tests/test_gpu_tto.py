"""Operator rounding tt_rounding(::TToperator) (src/tt_rounding.jl:275-277) on the
device: the rounding of the (n^2, R, R') reshape of the cores, against the oracle's
tto_rounding (itself pinned against the dense vectorised operator), and the
Hubbard sum H + H rounded back to 2H with H's bond dimension; tt_rounding! keeps a
non-orthogonal argument as it is (src/tt_rounding.jl:158-167)."""
import sys

import numpy as np
import pytest

import oracle as O
from helpers import to_api

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _api_op(T, A):
    return T.TToperator(A.N, [c.copy() for c in A.cores], A.dims, list(A.rks))


def _host_op(R):
    return O.TTO([np.asarray(c) for c in R.tto_vec], tuple(R.tto_dims), list(R.tto_rks))


def _direct_sum(H):
    L = H.N
    rk = [1] + [2 * r for r in H.rks[1:-1]] + [1]
    cores = []
    for k in range(L):
        c, R0, R1 = H.cores[k], H.rks[k], H.rks[k + 1]
        out = np.zeros((c.shape[0], c.shape[1], rk[k], rk[k + 1]))
        if k == 0:
            out[:, :, 0, :R1] = c[:, :, 0, :]
            out[:, :, 0, R1:] = c[:, :, 0, :]
        elif k == L - 1:
            out[:, :, :R0, 0] = c[:, :, :, 0]
            out[:, :, R0:, 0] = c[:, :, :, 0]
        else:
            out[:, :, :R0, :R1] = c
            out[:, :, R0:, R1:] = c
        cores.append(out)
    return O.TTO(cores, H.dims, rk)


def test_tto_rounding_hubbard_sum():
    import ttround as T
    say("case: tt_rounding(H + H), Hubbard L=5")
    H = O.hubbard_1d_mpo(5, 1.0, 4.0)
    HH = _direct_sum(H)
    R = _host_op(T.tt_rounding(_api_op(T, HH), tol=1e-12))
    ref = O.tto_rounding(HH, tol=1e-12)
    say(f"  ranks {R.rks} (oracle {ref.rks})")
    assert R.rks == ref.rks and max(R.rks) <= 6
    D = O.tto_to_tensor(R)
    np.testing.assert_allclose(D, 2 * O.hubbard_1d_dense(5, 1.0, 4.0), atol=1e-11)


@pytest.mark.parametrize("tol,rmax", [(1e-12, 2 ** 14), (0.2, 2 ** 14), (1e-12, 3)])
def test_tto_rounding_vs_oracle(tol, rmax):
    import ttround as T
    say(f"case: tt_rounding(TToperator) tol={tol} rmax={rmax}")
    g = np.random.default_rng(5)
    dims, rks = (2, 3, 2, 3), [1, 4, 6, 4, 1]
    A = O.TTO([g.standard_normal((n, n, rks[k], rks[k + 1])) for k, n in enumerate(dims)], dims, rks)
    R = _host_op(T.tt_rounding(_api_op(T, A), tol=tol, rmax=rmax))
    ref = O.tto_rounding(A, tol=tol, rmax=rmax)
    err = O.rel_err(O.tto_to_ttv(ref), O.tto_to_ttv(R))
    say(f"  ranks {R.rks}, rel err vs oracle {err:.3e}")
    assert R.rks == ref.rks and err < 1e-12


def test_tt_rounding_bang_keeps_non_orthogonal_input():
    import ttround as T
    x = O.rand_tt((2, 3, 2, 3), 3, np.random.default_rng(9))
    xa = to_api(x)
    before = [c.copy() for c in xa.ttv_vec]
    T.tt_rounding_(xa, tol=0.5)
    assert all(np.array_equal(a, b) for a, b in zip(xa.ttv_vec, before))
    y = T.orthogonalize(xa, 4)  # left-orthogonal: the rounding replaces it
    assert list(y.ttv_ot) == [1, 1, 1, 0]
    T.tt_rounding_(y, tol=0.5)
    assert list(y.ttv_ot) == [0, -1, -1, -1]
    ref = O.tt_rounding(O.orthogonalize(x, i=4), tol=0.5)
    assert list(y.ttv_rks) == ref.rks
