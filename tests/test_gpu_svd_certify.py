"""The SVD sweep's no-truncation certificate and its failure reporting.

At each bond the SVD sweep first tries a Cholesky certificate that the floor
rule (src/tt_rounding.jl:297-315) keeps every singular value; a certified bond
skips the SVD and keeps the QR/LQ factor as its orthogonal core.  These tests
build TTs whose singular values at one bond sit just below, at and above the
floor rule's threshold and check that the ranks equal the oracle's (which always
runs the SVD) and that the represented tensor agrees to 1e-12, with the
certificate on and off.  A handle whose Jacobi SVD is capped at one sweep must
report a numerical failure instead of returning an unconverged SVD
(src/tt_rounding.jl:191-195: the reference either converges or throws).
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


def _orth(rng, rows, cols):
    q, _ = np.linalg.qr(rng.standard_normal((rows, cols)))
    return q


def spectrum_tt(s, seed):
    """d = 4, n = 4 TT whose bond-2 singular values are exactly s (len 8): cores
    0-1 left-orthogonal, core 2 = diag(s) times a right-orthogonal core, core 3
    right-orthogonal."""
    rng = np.random.default_rng(seed)
    n, r1, r2, r3 = 4, 4, len(s), 4
    c0 = _orth(rng, n, r1).reshape(n, 1, r1, order="F")
    c1 = _orth(rng, n * r1, r2).reshape(n, r1, r2, order="F")  # left unfolding (n r1) x r2
    q2 = _orth(rng, n * r3, r2).T                               # right unfolding r2 x (n r3), orthonormal rows
    c2 = np.empty((n, r2, r3))
    for a in range(r2):
        row = s[a] * q2[a].reshape(n, r3, order="F")
        c2[:, a, :] = row
    q3 = _orth(rng, n, r3).T                                    # r3 x n, orthonormal rows
    c3 = np.empty((n, r3, 1))
    for a in range(r3):
        c3[:, a, 0] = q3[a]
    return O.TT([c0, c1, c2, c3], (n,) * 4, [1, r1, r2, r3, 1], [0, 0, 0, 0])


def handle(monkeypatch, **env):
    from ttround import _lib as L
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    h = L.Handle(0)
    for k in env:
        monkeypatch.delenv(k)
    return h


@pytest.mark.parametrize("c", [0.5, 0.99, 1.005, 1.2, 4.0])
def test_certificate_at_the_threshold(monkeypatch, c):
    """s = (1, ..., 1, x) with x^2 = c tol^2 sum(s^2): the floor rule drops x for
    c < 1 and keeps it for c >= 1; c = 1.005 lies inside the certificate's
    margin (the SVD decides), c = 1.2 and 4 are certified."""
    import ttround as T
    from ttround import _lib as L
    tol = 1e-3
    x2 = 7 * c * tol ** 2 / (1 - c * tol ** 2)
    s = np.array([1.0] * 7 + [np.sqrt(x2)])
    s[:7] *= np.linspace(1.0, 0.5, 7)  # distinct values, so no degenerate extension
    x2 = 1.0 * c * tol ** 2 * np.sum(s[:7] ** 2) / (1 - c * tol ** 2)
    s[7] = np.sqrt(x2)
    x = spectrum_tt(s, seed=int(1000 * c))
    ref = O.tt_rounding(x, tol=tol)
    say(f"case: certificate threshold c={c}: oracle ranks {ref.rks}")
    assert ref.rks[2] == (8 if c >= 1 else 7)
    res = {}
    for cert in ("1", "0"):
        h = handle(monkeypatch, TTR_SVD_CERTIFY=cert)
        monkeypatch.setattr(L, "_default", h)
        h.profile(True)
        y = to_oracle(T.tt_rounding(to_api(x), tol=tol))
        st = h.profile_stats()
        h.close()
        bonds = st.get("svd_certify:bonds", {"launches": 0, "flops": 0.0})
        say(f"  certify={cert}: ranks {y.rks}, bonds tried {bonds['launches']}, certified {int(bonds['flops'])}")
        assert y.rks == ref.rks and y.ot == ref.ot
        assert O.rel_err(ref, y) < 1e-12
        res[cert] = (y, bonds)
    # with the certificate every bond is tried; below the margin bond 2 is never certified
    assert res["1"][1]["launches"] == 3
    assert res["0"][1]["launches"] == 0
    if c >= 1.2:
        assert res["1"][1]["flops"] >= 1
    assert O.rel_err(res["0"][0], res["1"][0]) < 1e-12


def test_certificate_skipped_when_rmax_truncates(monkeypatch):
    """A bond of rank 8 with rmax 7 is truncated by rmax whatever its spectrum."""
    import ttround as T
    from ttround import _lib as L
    s = np.linspace(1.0, 0.5, 8)
    x = spectrum_tt(s, seed=7)
    ref = O.tt_rounding(x, tol=1e-3, rmax=7)
    h = handle(monkeypatch)
    monkeypatch.setattr(L, "_default", h)
    y = to_oracle(T.tt_rounding(to_api(x), tol=1e-3, rmax=7))
    h.close()
    say(f"case: rmax 7 truncation, ranks {y.rks}")
    assert y.rks == ref.rks and ref.rks[2] == 7
    assert O.rel_err(ref, y) < 1e-12


@pytest.mark.parametrize("direction", ["left", "right"])
def test_certified_and_svd_sweeps_agree(monkeypatch, direction):
    """Ranks 36..160, tolerance below every singular value: all bonds certified
    on one handle, all SVD on the other; both equal the oracle."""
    import ttround as T
    from ttround import _lib as L
    dims = (6,) * 6
    rng = np.random.default_rng(33)
    x = O.tt_add(O.rand_tt(dims, 100, rng), O.tt_scale(0.5, O.rand_tt(dims, 60, np.random.default_rng(34))))
    ref = O.tt_rounding(x, tol=1e-9, direction=direction)
    out = {}
    for cert in ("1", "0"):
        h = handle(monkeypatch, TTR_SVD_CERTIFY=cert)
        monkeypatch.setattr(L, "_default", h)
        h.profile(True)
        y = to_oracle(T.tt_rounding(to_api(x), tol=1e-9, direction=direction))
        st = h.profile_stats()
        h.close()
        bonds = st.get("svd_certify:bonds", {"launches": 0, "flops": 0.0})
        err = O.rel_err(ref, y)
        say(f"case: {direction} certify={cert}: ranks {y.rks}, certified {int(bonds['flops'])}/{bonds['launches']}, "
            f"rel err vs oracle {err:.3e}")
        assert y.rks == ref.rks and y.ot == ref.ot
        assert err < 1e-12
        out[cert] = (y, bonds)
    assert out["1"][1]["flops"] == out["1"][1]["launches"] == 5


def test_jacobi_nonconvergence_is_reported(monkeypatch):
    """One Jacobi sweep cannot diagonalise a random 36..160-column factor: the
    call fails with NumericalFailure; a normal handle then rounds the same TT."""
    import ttround as T
    from ttround import _lib as L
    dims = (6,) * 6
    x = O.tt_add(O.rand_tt(dims, 100, np.random.default_rng(33)),
                 O.tt_scale(1e-3, O.rand_tt(dims, 60, np.random.default_rng(34))))
    say("case: Jacobi capped at 1 sweep, certificate off")
    h = handle(monkeypatch, TTR_SVD_CERTIFY="0", TTR_JACOBI_MAX_SWEEPS="1")
    monkeypatch.setattr(L, "_default", h)
    with pytest.raises(L.NumericalFailure, match="did not converge"):
        T.tt_rounding(to_api(x), tol=1e-6)
    # the handle stays usable: a TT with no bond beyond one column block converges
    small = O.rand_tt((2, 2, 2), 1, np.random.default_rng(1))
    y = to_oracle(T.tt_rounding(to_api(small), tol=1e-12))
    assert O.rel_err(small, y) < 1e-12
    h.close()
    h = handle(monkeypatch)
    monkeypatch.setattr(L, "_default", h)
    y = to_oracle(T.tt_rounding(to_api(x), tol=1e-6))
    assert O.rel_err(O.tt_rounding(x, tol=1e-6), y) < 1e-12
    h.close()


def test_graded_terms_converge(monkeypatch):
    """Columns more than 1/eps apart in norm within one Jacobi visit (terms scaled by
    1e-10, 1e-20 and 1e-30): the SVD converges (the rotation threshold is also
    relative to the visit's largest column) and the rounding matches the oracle;
    the materialised 8-term sums of the 256-item benchmark are of this kind."""
    import ttround as T
    from ttround import _lib as L
    dims = (4,) * 8
    terms = [O.tt_scale(10.0 ** (-10 * j), O.rand_tt(dims, 12, np.random.default_rng(500 + j))) for j in range(4)]
    x = terms[0]
    for t in terms[1:]:
        x = O.tt_add(x, t)
    ref = O.tt_rounding(x, tol=1e-12)
    h = handle(monkeypatch, TTR_SVD_CERTIFY="0")
    monkeypatch.setattr(L, "_default", h)
    h.profile(True)
    y = to_oracle(T.tt_rounding(to_api(x), tol=1e-12))
    sw = h.profile_stats().get("jacobi_svd:sweeps", {})
    h.close()
    err = O.rel_err(ref, y)
    say(f"case: graded terms: ranks {y.rks} (oracle {ref.rks}), rel err {err:.3e}, "
        f"max sweeps {sw.get('bytes')}")
    assert y.rks == ref.rks and err < 1e-12
