"""FP64 numpy restatement of the TensorTrains rounding hot path (test oracle).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, by ``__graft_entry__.smoke()``
and by the ``cpu_baseline`` leg of ``bench.py`` as the checker; never by the
product package.

Conventions (mirroring the reference):
  * a TT core is a float64 ndarray indexed ``[i, a, b]`` with shape
    ``(n_k, r_{k-1}, r_k)`` (``src/tt_tools.jl:15,26,91``);
  * an operator core is indexed ``[i_out, j_in, a, b]`` with shape
    ``(n_k, n_k, R_{k-1}, R_k)`` (``src/tt_tools.jl:56,63``);
  * every Julia ``reshape`` is restated with ``order='F'`` (column-major);
  * indices in comments are Julia's 1-based ones, code is 0-based.

The random sketch draws are explicit inputs (``draws[k]`` holds the N(0,1)
numbers that ``randn!`` would put in the ``(bv_{k+1} n_k, bv_k, p_k)`` block
array, column-major), so the oracle and the device path see identical draws.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

__all__ = [
    "TT", "TTO", "jreshape", "r_and_d_to_rks", "zeros_tt", "rand_tt", "copy_tt",
    "ttv_to_tensor", "tto_to_tensor", "tt_add", "tt_scale", "tt_sub", "tt_lincomb",
    "mpo_apply", "tto_to_ttv", "ttv_to_tto", "is_leftorthogonal",
    "is_rightorthogonal", "orthogonalize", "tt_norm", "floor_rule", "tt_rounding",
    "tto_rounding", "default_rank_heuristic", "sketch_blocks_heuristic",
    "sketch_structure", "generate_sketch_blocks", "contract_sketch_core_backwards",
    "contract_sketch_core_backwards_op", "tt_recursive_sketch",
    "tt_recursive_sketch_op", "tt_sketch", "tt_combined_sketch", "sketch_structure_forward",
    "contract_sketch_core_forwards", "tt_recursive_sketch_forward", "contract_sketch_core_forwards_op",
    "tt_recursive_sketch_forward_op", "stable_inverse", "stta_sketch",
    "stta_sketch_tt", "stta_solve", "stta", "ttrand_rounding", "ttrand_rounding_lincomb",
    "ttrand_rounding_op", "ttrand_rounding_residual", "rel_err", "tt_dot", "dot",
    "dot_par", "partial_contraction", "dot_operator",
    "hubbard_1d_mpo", "hubbard_1d_dense", "hadamard", "contract_sketch_core_kronecker_backwards",
    "tt_recursive_sketch_tuple", "tt_sketch_tuple", "ttrand_rounding_tuple",
    "expand_basis", "sketched_rr", "sketched_rayleigh_ritz",
    "contract_sketch_core_kronecker_forwards", "tt_recursive_sketch_tuple_forward",
]


def jreshape(a: np.ndarray, shape) -> np.ndarray:
    """Julia ``reshape`` (column-major)."""
    return np.reshape(a, shape, order="F")


@dataclass
class TT:
    """``TTvector{Float64,N}`` (``src/tt_tools.jl:24-34``)."""
    cores: List[np.ndarray]
    dims: tuple
    rks: List[int]
    ot: List[int] = field(default_factory=list)

    def __post_init__(self):
        self.cores = [np.asarray(c, dtype=np.float64) for c in self.cores]
        self.dims = tuple(int(x) for x in self.dims)
        self.rks = [int(x) for x in self.rks]
        if not self.ot:
            self.ot = [0] * len(self.dims)
        self.ot = [int(x) for x in self.ot]

    @property
    def N(self) -> int:
        return len(self.dims)


@dataclass
class TTO:
    """``TToperator{Float64,N}`` (``src/tt_tools.jl:61-71``)."""
    cores: List[np.ndarray]
    dims: tuple
    rks: List[int]
    ot: List[int] = field(default_factory=list)

    def __post_init__(self):
        self.cores = [np.asarray(c, dtype=np.float64) for c in self.cores]
        self.dims = tuple(int(x) for x in self.dims)
        self.rks = [int(x) for x in self.rks]
        if not self.ot:
            self.ot = [0] * len(self.dims)

    @property
    def N(self) -> int:
        return len(self.dims)


# ---------------------------------------------------------------------------
# ranks, constructors, densification
# ---------------------------------------------------------------------------

def _wrapped_prod(vals) -> int:
    """Julia ``prod`` over ``Int64`` (wraps modulo 2^64)."""
    p = 1
    for v in vals:
        p = (p * int(v)) & 0xFFFFFFFFFFFFFFFF
    return p - (1 << 64) if p >= (1 << 63) else p


def r_and_d_to_rks(rks, dims, rmax: int = 1024) -> List[int]:
    """Cap each bond by the mode-size products (``src/tt_rounding.jl:6-24``).

    The reference's ``> 0`` tests guard against Int64 overflow of the
    products; the wrap-around is restated exactly."""
    d = len(dims)
    new = [1] * len(rks)
    for i in range(d):
        right = _wrapped_prod(dims[i:])
        left = _wrapped_prod(dims[:i])
        if right > 0:
            new[i] = min(rks[i], left, right, rmax) if left > 0 else min(rks[i], right, rmax)
        else:
            new[i] = min(rks[i], left, rmax) if left > 0 else min(rks[i], rmax)
    return [int(x) for x in new]


def zeros_tt(dims, rks, ot=None) -> TT:
    """``zeros_tt`` (``src/tt_tools.jl:89-93``)."""
    cores = [np.zeros((dims[i], rks[i], rks[i + 1])) for i in range(len(dims))]
    return TT(cores, tuple(dims), list(rks), list(ot) if ot is not None else [0] * len(dims))


def rand_tt(dims, rks, rng: np.random.Generator, stable: bool = True) -> TT:
    """N(0,1) cores as in ``rand_tt`` (``src/tt_tools.jl:255-318``, the
    non-orthogonal, non-normalised branch :300-303); ranks capped by
    ``r_and_d_to_rks(rks, dims; rmax=maximum(rks))`` (:264-265)."""
    if isinstance(rks, (int, np.integer)):
        r = int(rks)
        rks = [r] * (len(dims) + 1)
        rks = r_and_d_to_rks(rks, dims, rmax=r)
    elif stable:
        rks = r_and_d_to_rks(list(rks), dims, rmax=max(rks))
    cores = [rng.standard_normal((dims[i], rks[i], rks[i + 1])) for i in range(len(dims))]
    return TT(cores, tuple(dims), list(rks), [0] * len(dims))


def copy_tt(x: TT) -> TT:
    """``Base.copy`` (``src/tt_tools.jl:356-362``)."""
    return TT([c.copy() for c in x.cores], x.dims, list(x.rks), list(x.ot))


def ttv_to_tensor(x: TT) -> np.ndarray:
    """Dense tensor ``C[i_1..i_d] = A_1[i_1] ... A_d[i_d]`` (``src/tt_tools.jl:520-535``)."""
    acc = x.cores[0][:, 0, :]  # (n1, r1)
    for k in range(1, x.N):
        c = x.cores[k]  # (n, r, r')
        acc = np.tensordot(acc, c, axes=([acc.ndim - 1], [1]))  # (..., n, r')
    return acc[..., 0]


def tto_to_tensor(A: TTO) -> np.ndarray:
    """Dense operator as an ``(prod n) x (prod n)`` matrix with column-major
    multi-indices (index i_1 fastest), matching ``reshape(tto_to_tensor(A), n^d, :)``."""
    acc = A.cores[0][:, :, 0, :]  # (i1, j1, r1)
    for k in range(1, A.N):
        c = A.cores[k]
        acc = np.tensordot(acc, c, axes=([acc.ndim - 1], [2]))
    acc = acc[..., 0]
    d = A.N
    out_axes = [2 * k for k in range(d)]
    in_axes = [2 * k + 1 for k in range(d)]
    acc = np.transpose(acc, out_axes + in_axes)
    n = int(np.prod(A.dims))
    return jreshape(acc, (n, n))


# ---------------------------------------------------------------------------
# rank-growing operations
# ---------------------------------------------------------------------------

def tt_add(x: TT, y: TT) -> TT:
    """Block-diagonal direct sum (``src/tt_operations.jl:81-112``)."""
    assert x.dims == y.dims, "Incompatible dimensions"
    d = x.N
    rks = [a + b for a, b in zip(x.rks, y.rks)]
    rks[0] = 1
    rks[d] = 1
    cores = [np.zeros((x.dims[k], rks[k], rks[k + 1])) for k in range(d)]
    cores[0][:, :, : x.rks[1]] = x.cores[0]
    cores[0][:, :, x.rks[1]: rks[1]] = y.cores[0]
    for k in range(1, d - 1):
        cores[k][:, : x.rks[k], : x.rks[k + 1]] = x.cores[k]
        cores[k][:, x.rks[k]: rks[k], x.rks[k + 1]: rks[k + 1]] = y.cores[k]
    cores[d - 1][:, : x.rks[d - 1], 0] = x.cores[d - 1][:, :, 0]
    cores[d - 1][:, x.rks[d - 1]: rks[d - 1], 0] = y.cores[d - 1][:, :, 0]
    return TT(cores, x.dims, rks, [0] * d)


def tt_scale(a: float, x: TT) -> TT:
    """Scalar times TT: scales the first core with ``ot == 0`` only
    (``src/tt_operations.jl:428-441``)."""
    if a == 0:
        return zeros_tt(x.dims, [1] * (x.N + 1))
    if 0 not in x.ot:
        raise ValueError("no core with ot == 0 to scale")
    i = x.ot.index(0)
    cores = list(x.cores)
    cores[i] = a * cores[i]
    return TT(cores, x.dims, list(x.rks), list(x.ot))


def tt_sub(x: TT, y: TT) -> TT:
    """``x - y = (-1.0)*y + x`` (``src/tt_operations.jl:453-455``)."""
    return tt_add(tt_scale(-1.0, y), x)


def tt_lincomb(ys: Sequence[TT], coeffs: Sequence[float]) -> TT:
    """``*(A::Vector{TTvector}, x)`` (``src/tt_operations.jl:252-258``)."""
    out = tt_scale(coeffs[0], ys[0])
    for c, y in zip(coeffs[1:], ys[1:]):
        out = tt_add(out, tt_scale(c, y))
    return out


def mpo_apply(A: TTO, v: TT) -> TT:
    """``A*v`` with output bond index (operator rank fastest, then vector rank)
    (``src/tt_operations.jl:170-194``, index rule :180,184)."""
    assert A.dims == v.dims, "Incompatible dimensions"
    d = v.N
    rks = [a * b for a, b in zip(A.rks, v.rks)]
    cores = []
    for k in range(d):
        t = np.einsum("ijab,jcd->iacbd", A.cores[k], v.cores[k], optimize=True)
        n = v.dims[k]
        cores.append(jreshape(t, (n, rks[k], rks[k + 1])))
    return TT(cores, v.dims, rks, [0] * d)


def tto_to_ttv(A: TTO) -> TT:
    """``tto_to_ttv`` (``src/tt_tools.jl:603-611``)."""
    cores = [jreshape(A.cores[i], (A.dims[i] ** 2, A.rks[i], A.rks[i + 1])) for i in range(A.N)]
    return TT(cores, tuple(n * n for n in A.dims), list(A.rks), list(A.ot))


def ttv_to_tto(x: TT) -> TTO:
    """``ttv_to_tto`` (``src/tt_tools.jl:616-626``)."""
    nd = [int(math.isqrt(n)) for n in x.dims]
    assert all(m * m == n for m, n in zip(nd, x.dims)), "DimensionMismatch"
    cores = [jreshape(x.cores[i], (nd[i], nd[i], x.rks[i], x.rks[i + 1])) for i in range(x.N)]
    return TTO(cores, tuple(nd), list(x.rks), list(x.ot))


def tt_dot(x: TT, y: TT) -> float:
    """Euclidean inner product via the right-to-left partial contraction
    (``src/sketches.jl:1047-1070``)."""
    return float(partial_contraction(x, y, reverse=True)[0][0, 0])


def dot(x: TT, y: TT) -> float:
    """``dot(A, B)`` (``src/tt_operations.jl:302-321``): the left-to-right fold
    ``M[a,b] = A_k[z,α,a] B_k[z,β,b] out[α,β]`` starting from ``out = 1``."""
    assert x.dims == y.dims, "TT dimensions are not compatible"
    out = np.ones((1, 1))
    for k in range(x.N):
        out = np.einsum("zxa,zyb,xy->ab", x.cores[k], y.cores[k], out, optimize=True)
    return float(out[0, 0])


def dot_par(x: TT, y: TT) -> float:
    """``dot_par(A, B)`` (``src/tt_operations.jl:339-370``): per-core transfer
    matrices ``Y_k[(a,b),(c,d)] = A_k[z,a,c] B_k[z,b,d]`` multiplied from the
    right, ``C = Y_1 (Y_2 (... Y_d))``."""
    assert x.dims == y.dims, "TT dimensions are not compatible"
    Y = []
    for k in range(x.N):
        M = np.einsum("zac,zbd->abcd", x.cores[k], y.cores[k])
        Y.append(jreshape(M, (x.rks[k] * y.rks[k], x.rks[k + 1] * y.rks[k + 1])))
    C = Y[-1].reshape(-1, order="F")
    for k in range(x.N - 2, -1, -1):
        C = Y[k] @ C
    return float(C[0])


def partial_contraction(x: TT, y: TT, reverse: bool = True) -> List[np.ndarray]:
    """``partial_contraction(A, B; reverse)`` (``src/sketches.jl:1047-1070``):
    ``W[k]`` of shape ``(r^A_k, r^B_k)``; right to left
    ``W_k[a,b] = A_k[z,a,α] B_k[z,b,β] W_{k+1}[α,β]`` from ``W_{N+1} = 1``, left
    to right ``W_{k+1}[a,b] = A_k[z,α,a] B_k[z,β,b] W_k[α,β]`` from ``W_1 = 1``."""
    assert x.dims == y.dims, "TT dimensions are not compatible"
    d = x.N
    W = [np.zeros((x.rks[k], y.rks[k])) for k in range(d + 1)]
    if reverse:
        W[d] = np.ones((1, 1))
        for k in range(d - 1, -1, -1):
            W[k] = np.einsum("zax,zby,xy->ab", x.cores[k], y.cores[k], W[k + 1], optimize=True)
    else:
        W[0] = np.ones((1, 1))
        for k in range(d):
            W[k + 1] = np.einsum("zxa,zyb,xy->ab", x.cores[k], y.cores[k], W[k], optimize=True)
    return W


def dot_operator(psi: TT, H: TTO, phi: TT) -> float:
    """``dot_operator(psi, H, phi)`` = ``<psi|H|phi>`` (``src/tt_operations.jl:385-421``):
    pivot ``c`` = the later of the two orthogonality centres (first core with
    ``ot == 0``, else the last core); left environment over cores ``1..c`` when
    ``c > 1``, right environment over ``N..c+1`` (``N..1`` when ``c == 1``),
    combined by a full contraction."""
    N = psi.N
    assert psi.dims == phi.dims and tuple(H.dims) == tuple(psi.dims), "TT dimensions are not compatible"

    def center(v):
        for k, o in enumerate(v.ot):
            if o == 0:
                return k + 1
        return N

    c = max(center(psi), center(phi))
    # each environment update L[a,b,c] = L[x,y,z] A[i,x,a] H[i,j,y,b] B[j,z,c]
    # (and its mirror) as three pairwise BLAS contractions
    Lenv = np.ones((1, 1, 1))
    for k in range(c if c > 1 else 0):
        A, Hk, B = psi.cores[k], H.cores[k], phi.cores[k]
        T1 = np.tensordot(Lenv, B, axes=([2], [1]))          # (x, y, j, c)
        T2 = np.tensordot(T1, Hk, axes=([1, 2], [2, 1]))     # (x, c, i, b)
        Lenv = np.tensordot(T2, A, axes=([0, 2], [1, 0])).transpose(2, 1, 0)  # (a, b, c)
    Renv = np.ones((1, 1, 1))
    r = c + 1 if c > 1 else 1
    for k in range(N, r - 1, -1):
        A, Hk, B = psi.cores[k - 1], H.cores[k - 1], phi.cores[k - 1]
        T1 = np.tensordot(B, Renv, axes=([2], [2]))          # (j, z, a, b)
        T2 = np.tensordot(T1, Hk, axes=([0, 3], [1, 3]))     # (z, a, i, y)
        Renv = np.tensordot(A, T2, axes=([0, 2], [2, 1])).transpose(0, 2, 1)  # (x, y, z)
    if c == 1:
        return float(Renv[0, 0, 0])
    if c == N:
        return float(Lenv[0, 0, 0])
    return float(np.sum(Lenv * Renv))


# ---------------------------------------------------------------------------
# deterministic rounding
# ---------------------------------------------------------------------------

def is_leftorthogonal(x: TT) -> bool:
    """``src/tt_tools.jl:364-367``."""
    return all(o == 1 for o in x.ot[: x.N - 1])


def is_rightorthogonal(x: TT) -> bool:
    """``src/tt_tools.jl:369-372``."""
    return all(o == -1 for o in x.ot[1:])


def _lq(M: np.ndarray):
    """LQ through the QR of the transpose (LAPACK ``gelqf`` is ``geqrf`` on the
    rows, so the factors coincide)."""
    Q1, R1 = np.linalg.qr(M.T)
    return R1.T, Q1.T


def orthogonalize(x: TT, i: Optional[int] = None) -> TT:
    """Left/right orthogonalisation around root ``i`` (1-based, default 1)
    (``src/tt_rounding.jl:65-97``).  Left cores via Householder QR of the
    ``(r'_j n_j) x R_{j+1}`` unfolding with rows (alpha fastest, then i) (:72-79),
    right cores via LQ (:82-90), root ``FR * X_i[k] * FL`` (:93-95)."""
    d = x.N
    if i is None:
        i = 1
    assert 1 <= i <= d, "Impossible orthogonalization"
    y_rks = r_and_d_to_rks(x.rks, x.dims)
    cores: List[Optional[np.ndarray]] = [None] * d
    ot = [0] * d
    FR = np.ones((1, 1))
    for j in range(1, i):
        j0 = j - 1
        ot[j0] = 1
        n = x.dims[j0]
        t = np.einsum("ab,ibc->aic", FR, x.cores[j0], optimize=True)  # yleft[α_{j-1}, i_j, α_j]
        M = jreshape(t, (n * y_rks[j0], x.rks[j0 + 1]))
        Q, R = np.linalg.qr(M)
        y_rks[j0 + 1] = Q.shape[1]
        cores[j0] = np.transpose(jreshape(Q, (y_rks[j0], n, y_rks[j0 + 1])), (1, 0, 2))
        FR = R[: y_rks[j0 + 1], :]
    FL = np.ones((1, 1))
    for j in range(d, i, -1):
        j0 = j - 1
        ot[j0] = -1
        n = x.dims[j0]
        t = np.einsum("iab,bc->aci", x.cores[j0], FL, optimize=True)  # yright[α_{j-1}, β_j, i_j]
        M = jreshape(t, (x.rks[j0], y_rks[j0 + 1] * n))
        L, Q = _lq(M)
        y_rks[j0] = Q.shape[0]
        cores[j0] = np.transpose(jreshape(Q, (y_rks[j0], y_rks[j0 + 1], n)), (2, 0, 1))
        FL = L[:, : y_rks[j0]]
    i0 = i - 1
    ot[i0] = 0
    root = np.zeros((x.dims[i0], y_rks[i0], y_rks[i0 + 1]))
    for k in range(x.dims[i0]):
        root[k] = FR @ x.cores[i0][k] @ FL
    cores[i0] = root
    return TT(cores, x.dims, y_rks, ot)


def tt_norm(x: TT) -> float:
    """``norm(::TTvector)`` (``src/tt_rounding.jl:107-114``)."""
    zeros = [k for k, o in enumerate(x.ot) if o == 0]
    if len(zeros) == 1:
        return float(np.linalg.norm(x.cores[zeros[0]]))
    w = orthogonalize(x, i=x.N)
    return float(np.linalg.norm(w.cores[-1]))


def floor_rule(s: np.ndarray, tol: float, degen_tol: float = 1e-5):
    """Tail-energy truncation with degeneracy extension
    (``src/tt_rounding.jl:297-315``).  Returns ``(s[:k], k)``."""
    s = np.asarray(s, dtype=np.float64)
    d = len(s)
    if tol == 0.0:
        return s, d
    i = d
    weight = 0.0
    norm2 = float(np.dot(s, s))
    thr = tol ** 2 * norm2
    while i > 0 and weight < thr:
        weight += float(s[i - 1]) ** 2
        i -= 1
    i += 1
    while i < d:
        with np.errstate(divide="ignore", invalid="ignore"):
            ratio = np.float64(s[i]) / np.float64(s[i - 1])
        if np.isfinite(ratio) and abs(1.0 - ratio) <= degen_tol:
            i += 1
        else:
            break
    return s[:i], i


def _svd(M: np.ndarray):
    try:
        return np.linalg.svd(M, full_matrices=False)
    except np.linalg.LinAlgError:  # gesdd failure -> gesvd (tt_rounding.jl:191-195)
        import scipy.linalg
        return scipy.linalg.svd(M, full_matrices=False, lapack_driver="gesvd")


def _tt_rounding(y: TT, tol: float, rmax: int, direction: str) -> TT:
    """SVD truncation sweep on an orthogonalised TT (``src/tt_rounding.jl:174-269``)."""
    N = y.N
    dims = y.dims
    cores = list(y.cores)
    rks = list(y.rks)
    ot = list(y.ot)
    if direction == "left":
        if np.linalg.norm(cores[N - 1]) < tol:
            return zeros_tt(dims, rks)
        for j in range(N, 1, -1):
            j0 = j - 1
            nj, rprev, rj = dims[j0], rks[j0], rks[j0 + 1]
            M = jreshape(np.transpose(cores[j0], (1, 0, 2)), (rprev, nj * rj))
            u, s, vt = _svd(M)
            _, k = floor_rule(s, tol)
            k = min(k, rmax, int(np.sum(s > 0.0)))
            cores[j0] = np.transpose(jreshape(vt[:k, :], (k, nj, rj)), (1, 0, 2))
            US = u[:, :k] * s[:k][None, :]
            nl, rll = dims[j0 - 1], rks[j0 - 1]
            left = jreshape(cores[j0 - 1], (nl * rll, rprev))
            cores[j0 - 1] = jreshape(left @ US, (nl, rll, k))
            rks[j0] = k
            ot[j0] = -1
        ot[0] = 0
        return TT(cores, dims, rks, ot)
    elif direction == "right":
        if np.linalg.norm(cores[0]) < tol:
            return zeros_tt(dims, rks)
        for j in range(1, N):
            j0 = j - 1
            nj, rprev, rj = dims[j0], rks[j0], rks[j0 + 1]
            M = jreshape(cores[j0], (nj * rprev, rj))
            u, s, vt = _svd(M)
            _, k = floor_rule(s, tol)
            k = min(k, rmax, int(np.sum(s > 0.0)))
            cores[j0] = jreshape(u[:, :k], (nj, rprev, k))
            SVT = s[:k][:, None] * vt[:k, :]
            nn, rn = dims[j0 + 1], rks[j0 + 2]
            right = jreshape(np.transpose(cores[j0 + 1], (1, 0, 2)), (rj, nn * rn))
            cores[j0 + 1] = np.transpose(jreshape(SVT @ right, (k, nn, rn)), (1, 0, 2))
            rks[j0 + 1] = k
            ot[j0] = 1
        ot[N - 1] = 0
        return TT(cores, dims, rks, ot)
    raise ValueError("Invalid direction. Use direction=:right or direction=:left")


def tt_rounding(x: TT, tol: float = 1e-12, rmax: int = 2 ** 14, direction: str = "left") -> TT:
    """``tt_rounding`` (``src/tt_rounding.jl:146-154``)."""
    if direction == "left":
        y = copy_tt(x) if is_leftorthogonal(x) else orthogonalize(x, i=x.N)
    elif direction == "right":
        y = copy_tt(x) if is_rightorthogonal(x) else orthogonalize(x, i=1)
    else:
        raise ValueError("Invalid direction. Use direction=:right or direction=:left")
    return _tt_rounding(y, tol, rmax, direction)


def tto_rounding(A: TTO, tol: float = 1e-12, rmax: int = 2 ** 14) -> TTO:
    """``tt_rounding(::TToperator)`` (``src/tt_rounding.jl:275-277``)."""
    return ttv_to_tto(tt_rounding(tto_to_ttv(A), tol=tol, rmax=rmax))


# ---------------------------------------------------------------------------
# randomized rounding
# ---------------------------------------------------------------------------

def default_rank_heuristic(y) -> List[int]:
    """``src/tt_randtools.jl:31-50``: input ranks (element-wise max over a
    vector of TTs) plus ``ceil(0.5*max)`` on the interior bonds."""
    if isinstance(y, TT):
        rks = list(y.rks)
        N = y.N
    else:
        rks = list(y[0].rks)
        for t in y[1:]:
            rks = [max(a, b) for a, b in zip(rks, t.rks)]
        N = y[0].N
    ovs = int(math.ceil(0.5 * max(rks)))
    for k in range(1, N):
        rks[k] += ovs
    return rks


def sketch_blocks_heuristic(rks, bv, N: int) -> List[int]:
    """``compute_sketch_blocks_heuristic(...; reverse=true)`` (``src/sketches.jl:107-123``)."""
    p = [int(math.ceil(r / b)) for r, b in zip(rks, bv)]
    for i in range(1, N):
        p[i] = max(p[i], p[i - 1])
    p[N] = 1
    return p


def sketch_structure(dims, rks, block_rks: Optional[int] = None, p=None, oversampling=1):
    """Block ranks ``bv`` and block counts ``p`` of the reverse sketch
    (``src/sketches.jl:186-198``)."""
    N = len(dims)
    if block_rks is None:
        block_rks = N
    bv = [1] * (N + 1)
    for k in range(N):
        bv[k] = block_rks
    for k in range(N - 1, -1, -1):
        bv[k] = min(bv[k], dims[k] * bv[k + 1])
    if p is None or (isinstance(p, int) and p == 0):
        p = sketch_blocks_heuristic(rks, bv, N)
    else:
        p = [int(v) for v in p]
    if oversampling != 1:
        for k in range(N):
            p[k] = int(math.ceil(oversampling * p[k]))
    return bv, p


def generate_sketch_blocks(draws, left: int, dim: int, right: int, p: int, orthogonal: bool) -> np.ndarray:
    """Sketch blocks ``S (left, dim, right, p)`` from the N(0,1) draws in
    ``randn!`` order (``src/sketches.jl:44-84``): identity when orthogonal and
    ``right == dim*left`` (:46,58-65); per-block thin Householder Q scaled by
    ``sqrt(left*dim/right)`` when orthogonal (:71-78); ``1/sqrt(right)``
    otherwise (:80)."""
    assert (not orthogonal) or right <= dim * left
    use_identity = orthogonal and right == dim * left
    if use_identity:
        block = np.zeros((left * dim, right, p))
        for j in range(p):
            for i in range(right):
                block[i, i, j] = 1.0
    else:
        block = jreshape(np.asarray(draws, dtype=np.float64)[: left * dim * right * p], (left * dim, right, p)).copy()
        if orthogonal:
            for j in range(p):
                q, _ = np.linalg.qr(block[:, :, j])
                block[:, :, j] = q
            block *= math.sqrt(left * dim / right)
        else:
            block *= 1.0 / math.sqrt(right)
    return jreshape(block, (left, dim, right, p))


def _sketch_for(draws, k, left, dim, right, p, orthogonal):
    """``draws`` is either a list of per-core N(0,1) draw arrays (then the blocks
    are built by ``generate_sketch_blocks``) or a callable returning the final
    ``(left, dim, right, >= p)`` block array of core ``k`` (used to feed the
    oracle the exact blocks a device run used)."""
    if callable(draws):
        S = np.asarray(draws(k), dtype=np.float64)
        assert S.shape[:3] == (left, dim, right) and S.shape[3] >= p, (S.shape, (left, dim, right, p))
        return S[:, :, :, :p]
    return generate_sketch_blocks(draws[k], left, dim, right, p, orthogonal)


def sketch_uses_draws(left: int, dim: int, right: int, orthogonal: bool) -> bool:
    return not (orthogonal and right == dim * left)


def contract_sketch_core_backwards(A: np.ndarray, S: np.ndarray, V: np.ndarray) -> np.ndarray:
    """``W[a,b] = sum A[z,a,al] S[be,z,b] V[al,be]`` with the reference's
    cost-chosen order (``src/sketches.jl:1602-1638``)."""
    z, a, al = A.shape
    be, _, b = S.shape
    if a * be * (al + b + 1) < al * (b + 1) * (a + be):
        AV = jreshape(A, (z * a, al)) @ V  # (z*a, be)
        AV = jreshape(AV, (z, a, be))
        AVp = jreshape(np.transpose(AV, (1, 2, 0)), (a, be * z))
        return AVp @ jreshape(S, (be * z, b))
    VS = V @ jreshape(S, (be, z * b))  # (al, z*b)
    VS = jreshape(VS, (al * z, b))
    Ap = jreshape(np.transpose(A, (1, 2, 0)), (a, al * z))
    return Ap @ VS


def contract_sketch_core_backwards_op(A: np.ndarray, H: np.ndarray, S: np.ndarray, V: np.ndarray) -> np.ndarray:
    """``W[a,b,c] = sum A[z,a,al] H[ze,z,b,be] S[ga,ze,c] V[al,be,ga]`` with the
    reference's cost-chosen order, ((V A) H) S or ((V S) H) A
    (``src/sketches.jl:1648-1713``)."""
    z, a, al = A.shape
    ze, _, b, be = H.shape
    ga, _, c = S.shape
    if a * ga * (z * be * (al + ze * b) + ze * b * c) < al * c * (ze * be * (ga + z * b) + z * a * b):
        VA = jreshape(jreshape(A, (z * a, al)) @ jreshape(V, (al, be * ga)), (z, a, be, ga))
        VAp = jreshape(np.transpose(VA, (1, 3, 0, 2)), (a * ga, z * be))
        Bp = jreshape(np.transpose(H, (1, 3, 0, 2)), (z * be, ze * b))
        VAB = jreshape(VAp @ Bp, (a, ga, ze, b))
        VABp = jreshape(np.transpose(VAB, (0, 3, 1, 2)), (a * b, ga * ze))
        return jreshape(VABp @ jreshape(S, (ga * ze, c)), (a, b, c))
    VS = jreshape(jreshape(V, (al * be, ga)) @ jreshape(S, (ga, ze * c)), (al, be, ze, c))
    VSp = jreshape(np.transpose(VS, (2, 1, 0, 3)), (ze * be, al * c))
    Bp = jreshape(np.transpose(H, (1, 2, 0, 3)), (z * b, ze * be))
    VSB = jreshape(Bp @ VSp, (z, b, al, c))
    Ap = jreshape(np.transpose(A, (1, 0, 2)), (a, z * al))
    VSBp = jreshape(np.transpose(VSB, (0, 2, 1, 3)), (z * al, b * c))
    return jreshape(Ap @ VSBp, (a, b, c))


def _op_project(P: np.ndarray, y: np.ndarray, H: np.ndarray) -> np.ndarray:
    """``t[i,s,c,d] = sum_{a,b,j} P[s,a,b] y[j,a,c] H[i,j,b,d]`` as two products
    (the left-part update of ``src/tt_randtools.jl:337-339``)."""
    s_, a, b = P.shape
    j, _, c = y.shape
    i, _, _, d = H.shape
    U = jreshape(np.transpose(P, (0, 2, 1)), (s_ * b, a)) @ jreshape(np.transpose(y, (1, 0, 2)), (a, j * c))
    U = jreshape(U, (s_, b, j, c))
    Up = jreshape(np.transpose(U, (0, 3, 1, 2)), (s_ * c, b * j))
    Hp = jreshape(np.transpose(H, (2, 1, 0, 3)), (b * j, i * d))
    t = jreshape(Up @ Hp, (s_, c, i, d))
    return np.transpose(t, (2, 0, 1, 3))


def tt_recursive_sketch(A: TT, rks, draws, orthogonal: bool = True, block_rks: Optional[int] = None,
                        p=None, oversampling=1):
    """Reverse (right-to-left) recursive sketch (``src/sketches.jl:175-240,300-303``).

    Returns ``(W, sketch_rks)`` with ``W[k]`` of shape ``(A.rks[k], bv_k p_k)``
    (block index slowest), ``W[k]`` divided by ``sqrt(p_k)``."""
    N = A.N
    dims = A.dims
    assert rks[N] == 1 and A.rks[N] == 1
    bv, p = sketch_structure(dims, rks, block_rks, p, oversampling)
    W: List[Optional[np.ndarray]] = [None] * (N + 1)
    W[N] = np.ones((1, 1, 1))
    for k in range(N - 1, -1, -1):
        S = _sketch_for(draws, k, bv[k + 1], dims[k], bv[k], p[k], orthogonal)
        Wk = np.zeros((A.rks[k], bv[k], p[k]))
        for j in range(p[k]):
            Vn = W[k + 1][:, :, j if k < N - 1 else 0]
            Wk[:, :, j] = contract_sketch_core_backwards(A.cores[k], S[:, :, :, j], Vn)
        W[k] = Wk
    out = []
    for k in range(N + 1):
        W[k] = W[k] / math.sqrt(p[k])
        out.append(jreshape(W[k], (A.rks[k], bv[k] * p[k])))
    return out, [bv[k] * p[k] for k in range(N + 1)]


def tt_recursive_sketch_op(H: TTO, A: TT, rks, draws, orthogonal: bool = True,
                           block_rks: Optional[int] = None, p=None, oversampling=1):
    """Reverse sketch of ``H*A`` without forming it (``src/sketches.jl:317-389,456-459``).

    Returns ``W[k]`` of shape ``(A.rks[k], H.rks[k], s_k)``."""
    N = A.N
    dims = A.dims
    assert rks[N] == 1 and A.rks[N] == 1 and H.rks[N] == 1
    bv, p = sketch_structure(dims, rks, block_rks, p, oversampling)
    W: List[Optional[np.ndarray]] = [None] * (N + 1)
    W[N] = np.ones((1, 1, 1, 1))
    for k in range(N - 1, -1, -1):
        S = _sketch_for(draws, k, bv[k + 1], dims[k], bv[k], p[k], orthogonal)
        Wk = np.zeros((A.rks[k], H.rks[k], bv[k], p[k]))
        for j in range(p[k]):
            Vn = W[k + 1][:, :, :, j if k < N - 1 else 0]
            Wk[:, :, :, j] = contract_sketch_core_backwards_op(A.cores[k], H.cores[k], S[:, :, :, j], Vn)
        W[k] = Wk
    out = []
    for k in range(N + 1):
        W[k] = W[k] / math.sqrt(p[k])
        out.append(jreshape(W[k], (A.rks[k], H.rks[k], bv[k] * p[k])))
    return out, [bv[k] * p[k] for k in range(N + 1)]


def tt_sketch(A: TT, s: int, draws, orthogonal: bool = True, block_rks: Optional[int] = None,
              H: Optional[TTO] = None, reverse: bool = True):
    """Boundary-only reverse sketch ``tt_sketch(T, [H,] A, s; reverse=true)``
    (``src/sketches.jl:645-723``, operator form ``:729-813``).

    The same block count ``p = cld(s, bv_1)`` is used at every core and only
    the boundary ``W_1`` of shape ``(1, bv_1, p)`` (``(1, 1, bv_1, p)`` with an
    operator) is kept, divided by ``sqrt(p)`` once.  Returns
    ``(vec(W_1), bv_1 p)``.  ``reverse=False`` (``:679-694,796-811``): the forward
    sketch with ``p = cld(s, bv_{N+1})`` at every core, ``W_{N+1}`` kept; the same
    blocks as the forward recursive sketch with ``p = [1, p, ..., p]``."""
    N = A.N
    dims = A.dims
    if not reverse:
        bvf, _ = sketch_structure_forward(dims, [1] * (N + 1), block_rks)
        pc = -(-int(s) // bvf[N])
        pv = [1] + [pc] * N
        if H is None:
            W, _ = tt_recursive_sketch_forward(A, [1] * (N + 1), draws, orthogonal, block_rks, p=pv)
        else:
            W, _ = tt_recursive_sketch_forward_op(H, A, [1] * (N + 1), draws, orthogonal, block_rks, p=pv)
        return jreshape(W[N], -1), bvf[N] * pc
    bv, _ = sketch_structure(dims, [1] * (N + 1), block_rks)
    p = -(-int(s) // bv[0])
    W = np.ones((1, 1, 1)) if H is None else np.ones((1, 1, 1, 1))
    for k in range(N - 1, -1, -1):
        S = _sketch_for(draws, k, bv[k + 1], dims[k], bv[k], p, orthogonal)
        if H is None:
            Wk = np.zeros((A.rks[k], bv[k], p))
            for j in range(p):
                Wk[:, :, j] = contract_sketch_core_backwards(A.cores[k], S[:, :, :, j],
                                                             W[:, :, j if k < N - 1 else 0])
        else:
            Wk = np.zeros((A.rks[k], H.rks[k], bv[k], p))
            for j in range(p):
                Wk[:, :, :, j] = contract_sketch_core_backwards_op(A.cores[k], H.cores[k], S[:, :, :, j],
                                                                   W[:, :, :, j if k < N - 1 else 0])
        W = Wk
    return jreshape(W, -1) / math.sqrt(p), bv[0] * p


def tt_combined_sketch(A: TT, rks, s: int, draws, draws_extra, orthogonal: bool = True,
                       block_rks: Optional[int] = None, H: Optional[TTO] = None, reverse: bool = True):
    """``tt_combined_sketch([T,] [H,] A, rks, s; reverse=true)`` (``src/sketches.jl:930-953``,
    operator form ``:966-989``): the recursive sketch with the boundary rank raised
    to ``max(rks_1, min(s, rks_2))``; its boundary sketch, extended by a boundary-only
    sketch of size ``s - s_rec`` from the next seed (``draws_extra``) when ``s`` exceeds
    it, the two parts weighted by ``sqrt(len/s)``.  Returns ``(sketch, W, sketch_rks)``
    with ``W[1]`` replaced by ones and ``sketch_rks[1] = 1``.  ``reverse=False``: the
    same at the right boundary (``W[N+1]``, ``rks[N+1]`` raised to
    ``max(rks[N+1], min(s, rks[N]))``) with the forward sketches."""
    rks = list(rks)
    N = A.N
    b, inner = (0, 1) if reverse else (N, N - 1)
    rks[b] = max(rks[b], min(int(s), rks[inner]))
    if reverse:
        if H is None:
            W, srk = tt_recursive_sketch(A, rks, draws, orthogonal, block_rks)
        else:
            W, srk = tt_recursive_sketch_op(H, A, rks, draws, orthogonal, block_rks)
    else:
        if H is None:
            W, srk = tt_recursive_sketch_forward(A, rks, draws, orthogonal, block_rks)
        else:
            W, srk = tt_recursive_sketch_forward_op(H, A, rks, draws, orthogonal, block_rks)
    sketch = jreshape(W[b], -1)
    s_rec = srk[b]
    if s > s_rec:
        extra, _ = tt_sketch(A, s - s_rec, draws_extra, orthogonal, block_rks, H, reverse=reverse)
        s = s_rec + len(extra)
        sketch = np.concatenate([math.sqrt(len(sketch) / s) * sketch, math.sqrt(len(extra) / s) * extra])
    W[b] = np.ones((1, 1)) if H is None else np.ones((1, 1, 1))
    srk[b] = 1
    return sketch, W, srk


def sketch_structure_forward(dims, rks, block_rks: Optional[int] = None, p=None, oversampling=1):
    """Block ranks and counts of the forward (left-to-right) sketch
    (``src/sketches.jl:241-257``, ``compute_sketch_blocks_heuristic(...; reverse=false)``
    ``:113-121``): ``bv_1 = 1``, ``bv_{k+1} = min(block_rks, n_k bv_k)``; ``p_k =
    ceil(rks_k / bv_k)`` made non-increasing towards the right, ``p_1 = 1``; the
    oversampling multiplies ``p_2..p_{N+1}``."""
    N = len(dims)
    if block_rks is None:
        block_rks = N
    bv = [1] * (N + 1)
    for k in range(1, N + 1):
        bv[k] = block_rks
    for k in range(N):
        bv[k + 1] = min(bv[k + 1], dims[k] * bv[k])
    if p is None or (isinstance(p, int) and p == 0):
        p = [int(math.ceil(r / b)) for r, b in zip(rks, bv)]
        for i in range(N - 1, 0, -1):
            p[i] = max(p[i], p[i + 1])
        p[0] = 1
    else:
        p = [int(v) for v in p]
    if oversampling != 1:
        for k in range(1, N + 1):
            p[k] = int(math.ceil(oversampling * p[k]))
    return bv, p


def contract_sketch_core_forwards(A: np.ndarray, S: np.ndarray, V: np.ndarray) -> np.ndarray:
    """``W[a,b] = sum A[z,al,a] S[be,z,b] V[al,be]`` with the reference's
    cost-chosen order, ((V^T A) S) or ((V S) A) (``src/sketches.jl:1192-1230``)."""
    z, al, a = A.shape
    be, _, b = S.shape
    if a * (be + 1) * (al + b) < al * b * (a + be + 1):
        VA = V.T @ jreshape(np.transpose(A, (1, 0, 2)), (al, z * a))  # (be, z*a)
        return jreshape(VA, (be * z, a)).T @ jreshape(S, (be * z, b))
    VS = jreshape(V @ jreshape(S, (be, z * b)), (al, z, b))
    VSp = jreshape(np.transpose(VS, (1, 0, 2)), (z * al, b))
    return jreshape(A, (z * al, a)).T @ VSp


def tt_recursive_sketch_forward(A: TT, rks, draws, orthogonal: bool = True, block_rks: Optional[int] = None,
                                p=None, oversampling=1):
    """Forward (left-to-right) recursive sketch, ``tt_recursive_sketch(...; reverse=false)``
    (``src/sketches.jl:241-298,300-303``).  Core ``k`` uses blocks
    ``(bv_k, n_k, bv_{k+1}, p_{k+1})``; ``W[k+1][:,:,j]`` contracts core ``k``
    with block ``j`` and ``W[k][:,:,j]`` (``W[1] = 1``).  Returns ``W[k]`` of
    shape ``(A.rks[k], bv_k p_k)`` divided by ``sqrt(p_k)``, and the sketch ranks."""
    N = A.N
    dims = A.dims
    assert rks[0] == 1 and A.rks[0] == 1
    bv, p = sketch_structure_forward(dims, rks, block_rks, p, oversampling)
    W: List[Optional[np.ndarray]] = [None] * (N + 1)
    W[0] = np.ones((1, 1, 1))
    for k in range(N):
        S = _sketch_for(draws, k, bv[k], dims[k], bv[k + 1], p[k + 1], orthogonal)
        Wk = np.zeros((A.rks[k + 1], bv[k + 1], p[k + 1]))
        for j in range(p[k + 1]):
            Wk[:, :, j] = contract_sketch_core_forwards(A.cores[k], S[:, :, :, j], W[k][:, :, j if k > 0 else 0])
        W[k + 1] = Wk
    out = []
    for k in range(N + 1):
        out.append(jreshape(W[k] / math.sqrt(p[k]), (A.rks[k], bv[k] * p[k])))
    return out, [bv[k] * p[k] for k in range(N + 1)]


def contract_sketch_core_forwards_op(A: np.ndarray, H: np.ndarray, S: np.ndarray, V: np.ndarray) -> np.ndarray:
    """``W[a,b,c] = sum A[ze,al,a] H[z,ze,be,b] S[ga,z,c] V[al,be,ga]`` (the forward
    operator contraction, ``src/sketches.jl:1240-1309``) with the reference's
    cost-chosen order: ((V A) H) S or ((V S) H) A."""
    ze, al, a = A.shape
    z, _, be, b = H.shape
    ga, _, c = S.shape
    if ze * a * al * be * ga + a * ga * ze * be * z * b + a * b * ga * z * c < \
            al * be * ga * z * c + ze * b * z * be * al * c + a * ze * al * b * c:
        Ap = jreshape(np.transpose(A, (0, 2, 1)), (ze * a, al))                       # [ze, a, al]
        VA = jreshape(Ap @ jreshape(V, (al, be * ga)), (ze, a, be, ga))
        VAp = jreshape(np.transpose(VA, (1, 3, 0, 2)), (a * ga, ze * be))            # [a, ga, ze, be]
        Bp = jreshape(np.transpose(H, (1, 2, 0, 3)), (ze * be, z * b))                # [ze, be, z, b]
        VAB = jreshape(VAp @ Bp, (a, ga, z, b))
        VABp = jreshape(np.transpose(VAB, (0, 3, 1, 2)), (a * b, ga * z))             # [a, b, ga, z]
        return jreshape(VABp @ jreshape(S, (ga * z, c)), (a, b, c))
    VS = jreshape(jreshape(V, (al * be, ga)) @ jreshape(S, (ga, z * c)), (al, be, z, c))
    VSp = jreshape(np.transpose(VS, (2, 1, 0, 3)), (z * be, al * c))                  # [z, be, al, c]
    Bp = jreshape(np.transpose(H, (1, 3, 0, 2)), (ze * b, z * be))                    # [ze, b, z, be]
    VBS = jreshape(Bp @ VSp, (ze, b, al, c))
    Ap = jreshape(np.transpose(A, (2, 0, 1)), (a, ze * al))                           # [a, ze, al]
    VBSp = jreshape(np.transpose(VBS, (0, 2, 1, 3)), (ze * al, b * c))               # [ze, al, b, c]
    return jreshape(Ap @ VBSp, (a, b, c))


def tt_recursive_sketch_forward_op(H: TTO, A: TT, rks, draws, orthogonal: bool = True,
                                   block_rks: Optional[int] = None, p=None, oversampling=1):
    """Forward recursive sketch of ``H A`` without forming it,
    ``tt_recursive_sketch(T, H, A, rks; reverse=false)`` (``src/sketches.jl:390-459``).
    The block structure is the forward one, and, as in the reference, no
    oversampling is applied on this path (``:399-402``).  Returns ``W[k]`` of shape
    ``(A.rks[k], H.rks[k], bv_k p_k)`` divided by ``sqrt(p_k)`` and the sketch ranks."""
    N = A.N
    dims = A.dims
    assert rks[0] == 1 and A.rks[0] == 1 and H.rks[0] == 1
    bv, p = sketch_structure_forward(dims, rks, block_rks, p, 1)
    W: List[Optional[np.ndarray]] = [None] * (N + 1)
    W[0] = np.ones((1, 1, 1, 1))
    for k in range(N):
        S = _sketch_for(draws, k, bv[k], dims[k], bv[k + 1], p[k + 1], orthogonal)
        Wk = np.zeros((A.rks[k + 1], H.rks[k + 1], bv[k + 1], p[k + 1]))
        for j in range(p[k + 1]):
            Wk[:, :, :, j] = contract_sketch_core_forwards_op(A.cores[k], H.cores[k], S[:, :, :, j],
                                                              W[k][:, :, :, j if k > 0 else 0])
        W[k + 1] = Wk
    out = [jreshape(W[k] / math.sqrt(p[k]), (A.rks[k], H.rks[k], bv[k] * p[k])) for k in range(N + 1)]
    return out, [bv[k] * p[k] for k in range(N + 1)]


def stable_inverse(A: np.ndarray, eps: float = 1e-12) -> np.ndarray:
    """Pseudo-inverse over the singular values above ``eps * max(s)``
    (``src/tt_randtools.jl:485-488``)."""
    u, s, vt = np.linalg.svd(A, full_matrices=False)
    keep = s > s.max() * eps
    return vt[keep].T @ np.diag(1.0 / s[keep]) @ u[:, keep].T


def stta_sketch(x: TT, rks, draws_left, draws_right, orthogonal: bool = True, block_rks: Optional[int] = None):
    """``stta_sketch(x, rks; seed_left, seed_right, orthogonal, block_rks)``
    (``src/sketches.jl:1126-1156``): the forward sketch with oversampling 1.5 and
    the reverse sketch, ``Omega[k] = W_L[k+1]^T W_R[k+1]`` (k = 1..N-1) and
    ``Psi[k][i,a,b] = x_k[i,al,be] W_L[k][al,a] W_R[k+1][be,b]`` (k = 1..N)."""
    N = x.N
    WL, sl = tt_recursive_sketch_forward(x, rks, draws_left, orthogonal, block_rks, oversampling=1.5)
    WR, sr = tt_recursive_sketch(x, rks, draws_right, orthogonal, block_rks)
    Om = [WL[k + 1].T @ WR[k + 1] for k in range(N - 1)]
    Ps = [np.einsum("iab,ac,bd->icd", x.cores[k], WL[k], WR[k + 1], optimize=True) for k in range(N)]
    return Om, Ps, sl, sr


def stta_sketch_tt(x: TT, L: TT, R: TT):
    """``stta_sketch(x, L, R)`` with explicit sketch TTs (``src/sketches.jl:1112-1124``):
    the left and right interface contractions of ``x`` with ``L`` and ``R``."""
    lc = partial_contraction(x, L, reverse=False)
    rc = partial_contraction(x, R, reverse=True)
    N = x.N
    Om = [lc[k + 1].T @ rc[k + 1] for k in range(N - 1)]
    Ps = [np.einsum("iyz,ya,zb->iab", x.cores[k], lc[k], rc[k + 1], optimize=True) for k in range(N)]
    return Om, Ps


def stta_solve(Om, Ps, dims) -> TT:
    """The local solves of ``stta`` (``src/tt_randtools.jl:562-575``): per bond,
    ``Psi[k] pinv(Omega[k])`` when ``Omega[k]`` has fewer rows than columns,
    else ``pinv(Omega[k]) Psi[k+1]`` (in the second index of ``Psi[k+1]``)."""
    N = len(dims)
    P = [c.copy() for c in Ps]
    rks = [1] * (N + 1)
    for k in range(N - 1):
        O_ = Om[k]
        if O_.shape[0] < O_.shape[1]:
            n_, a, b = P[k].shape
            P[k] = jreshape(jreshape(P[k], (n_ * a, b)) @ stable_inverse(O_), (n_, a, -1))
            rks[k + 1] = P[k].shape[2]
        else:
            n_, a, b = P[k + 1].shape
            t = stable_inverse(O_) @ jreshape(np.transpose(P[k + 1], (1, 0, 2)), (a, n_ * b))
            P[k + 1] = np.transpose(jreshape(t, (-1, n_, b)), (1, 0, 2))
            rks[k + 1] = P[k + 1].shape[1]
    return TT(P, dims, rks, [0] * N)


def stta(alpha, ys: Sequence[TT], rks, draws_left, draws_right, orthogonal: bool = True,
         block_rks: Optional[int] = None) -> TT:
    """``stta(alpha, y, rks; seed_left, seed_right, orthogonal, block_rks)``
    (``src/tt_randtools.jl:552-638``; one term is ``stta(y, rks)``): the sketches
    of ``sum_j alpha_j y_j`` (linear in the terms), then the local solves."""
    Om, Ps = None, None
    for a, y in zip(alpha, ys):
        O_, P_, _, _ = stta_sketch(y, rks, draws_left, draws_right, orthogonal, block_rks)
        if Om is None:
            Om, Ps = [a * o for o in O_], [a * q for q in P_]
        else:
            Om = [o + a * o2 for o, o2 in zip(Om, O_)]
            Ps = [q + a * q2 for q, q2 in zip(Ps, P_)]
    return stta_solve(Om, Ps, ys[0].dims)


# ------------------------------------------------------- entrywise (Hadamard) products

def hadamard(xs: Sequence[TT], first_fastest: bool = False) -> TT:
    """Entrywise product of TTs: slice ``i`` of core ``k`` is the Kronecker product
    of the factors' slices.  ``first_fastest=False``: ``x * y`` of
    ``src/tt_operations.jl:273-290`` (``kron(x[i,:,:], y[i,:,:])``, so the last
    factor's rank index runs fastest), ot all zero; ``True``: the first factor's
    index fastest, the order of the NTuple methods (``src/tt_randtools.jl:378-382``,
    ``src/sketches.jl:574-575``)."""
    xs = list(xs)
    dims, N = xs[0].dims, xs[0].N
    assert all(t.dims == dims for t in xs)
    order = xs if not first_fastest else xs[::-1]  # np.kron: the later operand fastest
    cores = []
    for k in range(N):
        c = order[0].cores[k]
        for t in order[1:]:
            c = np.stack([np.kron(c[i], t.cores[k][i]) for i in range(dims[k])])
        cores.append(c)
    rks = [int(np.prod([t.rks[k] for t in xs])) for k in range(N + 1)]
    return TT(cores, dims, rks, [0] * N)


def contract_sketch_core_kronecker_backwards(As: Sequence[np.ndarray], S: np.ndarray, V: np.ndarray) -> np.ndarray:
    """``W[a_1..a_M, b] = sum_{z, al, be} prod_i A_i[z, a_i, al_i] S[be, z, b]
    V[al_1..al_M, be]`` (``src/sketches.jl:1735-1881``), factor by factor per mode
    index, the product core never formed."""
    M = len(As)
    out = None
    for z in range(As[0].shape[0]):
        T = V
        for i in range(M):
            T = np.moveaxis(np.tensordot(As[i][z], T, axes=([1], [i])), 0, i)
        T = np.tensordot(T, S[:, z, :], axes=([M], [0]))
        out = T if out is None else out + T
    return out


def tt_recursive_sketch_tuple(ys: Sequence[TT], rks, draws, orthogonal: bool = True,
                              block_rks: Optional[int] = None, p=None, oversampling=1):
    """Reverse sketch of the entrywise product of ``ys`` without forming it
    (``src/sketches.jl:474-535,600-604``).  ``W[k]`` has shape
    ``(r_1k, ..., r_Mk, bv_k p_k)``."""
    ys = list(ys)
    M, N, dims = len(ys), ys[0].N, ys[0].dims
    assert rks[N] == 1 and all(t.rks[N] == 1 for t in ys)
    bv, p = sketch_structure(dims, rks, block_rks, p, oversampling)
    W: List[Optional[np.ndarray]] = [None] * (N + 1)
    W[N] = np.ones((1,) * (M + 2))
    for k in range(N - 1, -1, -1):
        S = _sketch_for(draws, k, bv[k + 1], dims[k], bv[k], p[k], orthogonal)
        Wk = np.zeros(tuple(t.rks[k] for t in ys) + (bv[k], p[k]))
        for j in range(p[k]):
            Vn = W[k + 1][..., j if k < N - 1 else 0]
            Wk[..., j] = contract_sketch_core_kronecker_backwards([t.cores[k] for t in ys], S[:, :, :, j], Vn)
        W[k] = Wk
    out = []
    for k in range(N + 1):
        out.append(jreshape(W[k] / math.sqrt(p[k]), tuple(t.rks[k] for t in ys) + (bv[k] * p[k],)))
    return out, [bv[k] * p[k] for k in range(N + 1)]


def contract_sketch_core_kronecker_forwards(As: Sequence[np.ndarray], S: np.ndarray, V: np.ndarray) -> np.ndarray:
    """``W[a_1..a_M, b] = sum_{z, al, be} prod_i A_i[z, al_i, a_i] S[be, z, b]
    V[al_1..al_M, be]`` (``src/sketches.jl:1332-1478``), factor by factor per mode
    index, the product core never formed."""
    M = len(As)
    out = None
    for z in range(As[0].shape[0]):
        T = V
        for i in range(M):
            T = np.moveaxis(np.tensordot(As[i][z], T, axes=([0], [i])), 0, i)
        T = np.tensordot(T, S[:, z, :], axes=([M], [0]))
        out = T if out is None else out + T
    return out


def tt_recursive_sketch_tuple_forward(ys: Sequence[TT], rks, draws, orthogonal: bool = True,
                                      block_rks: Optional[int] = None, p=None, oversampling=1):
    """Forward (left-to-right) sketch of the entrywise product of ``ys`` without
    forming it (``src/sketches.jl:539-604``): ``W[k+1][..., j]`` contracts the
    factors' cores ``k`` and block ``j`` with ``W[k][..., j]`` (``W[1] = 1``).
    ``W[k]`` has shape ``(r_1k, ..., r_Mk, bv_k p_k)``, divided by ``sqrt(p_k)``."""
    ys = list(ys)
    M, N, dims = len(ys), ys[0].N, ys[0].dims
    assert rks[0] == 1 and all(t.rks[0] == 1 for t in ys)
    bv, p = sketch_structure_forward(dims, rks, block_rks, p, oversampling)
    W: List[Optional[np.ndarray]] = [None] * (N + 1)
    W[0] = np.ones((1,) * (M + 2))
    for k in range(N):
        S = _sketch_for(draws, k, bv[k], dims[k], bv[k + 1], p[k + 1], orthogonal)
        Wk = np.zeros(tuple(t.rks[k + 1] for t in ys) + (bv[k + 1], p[k + 1]))
        for j in range(p[k + 1]):
            Vp = W[k][..., j if k > 0 else 0]
            Wk[..., j] = contract_sketch_core_kronecker_forwards([t.cores[k] for t in ys], S[:, :, :, j], Vp)
        W[k + 1] = Wk
    out = []
    for k in range(N + 1):
        out.append(jreshape(W[k] / math.sqrt(p[k]), tuple(t.rks[k] for t in ys) + (bv[k] * p[k],)))
    return out, [bv[k] * p[k] for k in range(N + 1)]


def tt_sketch_tuple(ys: Sequence[TT], s: int, draws, orthogonal: bool = True, block_rks: Optional[int] = None,
                    reverse: bool = True):
    """Boundary sketch of the entrywise product (``src/sketches.jl:819-909``):
    the recursive sketch with ``p = cld(s, bv_boundary)`` blocks at every core
    (boundary ``W_1`` right to left, ``W_{N+1}`` left to right)."""
    ys = list(ys)
    N, dims = ys[0].N, ys[0].dims
    if reverse:
        bv, _ = sketch_structure(dims, [1] * (N + 1), block_rks)
        pc = -(-int(s) // bv[0])
        W, srk = tt_recursive_sketch_tuple(ys, [1] * (N + 1), draws, orthogonal, block_rks, p=[pc] * N + [1])
        return np.reshape(W[0], -1, order="F"), bv[0] * pc
    bv, _ = sketch_structure_forward(dims, [1] * (N + 1), block_rks)
    pc = -(-int(s) // bv[N])
    W, srk = tt_recursive_sketch_tuple_forward(ys, [1] * (N + 1), draws, orthogonal, block_rks, p=[1] + [pc] * N)
    return np.reshape(W[N], -1, order="F"), bv[N] * pc


def ttrand_rounding_tuple(ys: Sequence[TT], rks, draws, orthogonal: bool = True,
                          block_rks: Optional[int] = None) -> TT:
    """Randomized rounding of the entrywise product of ``ys``, never formed
    (``src/tt_randtools.jl:364-458``): the left part ``y_k`` carries the factors'
    rank indices (first fastest) and is advanced one factor at a time."""
    ys = list(ys)
    M, N, dims = len(ys), ys[0].N, ys[0].dims
    W, srk = tt_recursive_sketch_tuple(ys, rks, draws, orthogonal, block_rks)
    vec: List[Optional[np.ndarray]] = [None] * N
    out_rks = [1] * (N + 1)
    ot = [0] * N
    # y_1[i, 1, al_1..al_M] = prod_f y_f[i, 1, al_f]
    Y = np.ones((dims[0], 1) + (1,) * M)
    for f, t in enumerate(ys):
        shp = [dims[0], 1] + [1] * M
        shp[2 + f] = t.rks[1]
        Y = Y * t.cores[0].reshape(shp)
    for k in range(N - 1):
        R = [t.rks[k + 1] for t in ys]
        Yf = jreshape(Y, (dims[k] * out_rks[k], int(np.prod(R))))
        Z = Yf @ jreshape(W[k + 1], (int(np.prod(R)), srk[k + 1]))
        Q = _qr_thin(Z)
        out_rks[k + 1] = Q.shape[1]
        vec[k] = jreshape(Q, (dims[k], out_rks[k], out_rks[k + 1]))
        ot[k] = 1
        # Qy[al_1..al_M, rho] = sum_{i, r} y_k[i, r, al] Q[i, r, rho], then each factor's
        # next core applied to its own index
        T = jreshape(Q.T @ Yf, (out_rks[k + 1],) + tuple(R))  # (rho, al_1..al_M)
        n1 = dims[k + 1]
        T = np.broadcast_to(T[None], (n1,) + T.shape).copy()  # (i, rho, al_1..al_M)
        for f, t in enumerate(ys):
            # contract al_f (axis 2 + f) with core[i, al_f, be_f] for every i
            T = np.einsum(T, list(range(M + 2)), t.cores[k + 1], [0, 2 + f, M + 2],
                          [x if x != 2 + f else M + 2 for x in range(M + 2)], optimize=True)
        Y = T
    vec[N - 1] = jreshape(Y, (dims[N - 1], out_rks[N - 1], 1))
    return TT(vec, dims, out_rks, ot)


def _qr_thin(Z: np.ndarray) -> np.ndarray:
    Q, _ = np.linalg.qr(Z)
    return Q


def ttrand_rounding(y: TT, rks, draws, orthogonal: bool = True, block_rks: Optional[int] = None) -> TT:
    """Randomize-then-orthogonalize rounding of one TT (``src/tt_randtools.jl:112-154``)."""
    return ttrand_rounding_lincomb([1.0], [y], rks, draws, orthogonal, block_rks, _single=True)


def ttrand_rounding_lincomb(alpha, ys: Sequence[TT], rks, draws, orthogonal: bool = True,
                            block_rks: Optional[int] = None, _single: bool = False) -> TT:
    """Rounding of ``sum_j alpha_j y_j`` with one sketch per term drawn from the
    same draws (``src/tt_randtools.jl:223-276``)."""
    m = len(ys)
    dims = ys[0].dims
    N = ys[0].N
    assert all(t.dims == dims for t in ys)
    Ws = []
    srk = None
    for t in ys:
        W, s = tt_recursive_sketch(t, rks, draws, orthogonal, block_rks)
        assert srk is None or s == srk
        srk = s
        Ws.append(W)
    vec: List[Optional[np.ndarray]] = [None] * N
    out_rks = [1] * (N + 1)
    ot = [0] * N
    if _single:
        Y = [jreshape(ys[0].cores[0], (dims[0], 1, ys[0].rks[1]))]
    else:
        Y = [alpha[j] * jreshape(ys[j].cores[0], (dims[0], 1, ys[j].rks[1])) for j in range(m)]
    for k in range(N - 1):
        Z = np.zeros((dims[k], out_rks[k], srk[k + 1]))
        for j in range(m):
            Z += np.tensordot(Y[j], Ws[j][k + 1], axes=([2], [0]))
        Z = jreshape(Z, (dims[k] * out_rks[k], srk[k + 1]))
        Q = _qr_thin(Z)
        out_rks[k + 1] = Q.shape[1]
        vec[k] = jreshape(Q, (dims[k], out_rks[k], out_rks[k + 1]))
        ot[k] = 1
        for j in range(m):
            P = np.einsum("ira,irs->sa", Y[j], vec[k], optimize=True)
            Y[j] = np.einsum("sa,iab->isb", P, ys[j].cores[k + 1], optimize=True)
    out_rks[N] = 1
    last = Y[0].copy()
    for j in range(1, m):
        last = last + Y[j]
    vec[N - 1] = jreshape(last, (dims[N - 1], out_rks[N - 1], out_rks[N]))
    return TT(vec, dims, out_rks, ot)


def ttrand_rounding_op(alpha, H: TTO, ys: Sequence[TT], rks, draws, orthogonal: bool = True,
                       block_rks: Optional[int] = None) -> TT:
    """Rounding of ``alpha_1 H y_1 + sum_{j>1} alpha_j y_j`` with the implicit
    operator sketch (``src/tt_randtools.jl:285-354``)."""
    m = len(ys)
    dims = ys[0].dims
    N = ys[0].N
    W1, srk = tt_recursive_sketch_op(H, ys[0], rks, draws, orthogonal, block_rks)
    Ws = [[jreshape(W1[k], (ys[0].rks[k] * H.rks[k], srk[k])) for k in range(N + 1)]]
    for t in ys[1:]:
        W, s = tt_recursive_sketch(t, rks, draws, orthogonal, block_rks)
        assert s == srk
        Ws.append(W)
    vec: List[Optional[np.ndarray]] = [None] * N
    out_rks = [1] * (N + 1)
    ot = [0] * N
    Ay1 = np.einsum("jab,ijcd->iacbd", ys[0].cores[0], H.cores[0], optimize=True)  # (n,1,1,Ry2,RA2)
    Ay1 = alpha[0] * Ay1
    Y = [jreshape(Ay1, (dims[0], 1, ys[0].rks[1] * H.rks[1]))]
    for j in range(1, m):
        Y.append(alpha[j] * jreshape(ys[j].cores[0], (dims[0], 1, ys[j].rks[1])))
    for k in range(N - 1):
        Z = np.zeros((dims[k], out_rks[k], srk[k + 1]))
        for j in range(m):
            Z += np.tensordot(Y[j], Ws[j][k + 1], axes=([2], [0]))
        Z = jreshape(Z, (dims[k] * out_rks[k], srk[k + 1]))
        Q = _qr_thin(Z)
        out_rks[k + 1] = Q.shape[1]
        vec[k] = jreshape(Q, (dims[k], out_rks[k], out_rks[k + 1]))
        ot[k] = 1
        y1k = jreshape(Y[0], (dims[k], out_rks[k], ys[0].rks[k + 1], H.rks[k + 1]))
        P = np.einsum("irab,irs->sab", y1k, vec[k], optimize=True)  # (rho', alpha, beta)
        t = _op_project(P, ys[0].cores[k + 1], H.cores[k + 1])
        Ynew = [jreshape(t, (dims[k + 1], out_rks[k + 1], ys[0].rks[k + 2] * H.rks[k + 2]))]
        for j in range(1, m):
            Pj = np.einsum("ira,irs->sa", Y[j], vec[k], optimize=True)
            Ynew.append(np.einsum("sa,iab->isb", Pj, ys[j].cores[k + 1], optimize=True))
        Y = Ynew
    out_rks[N] = 1
    last = Y[0].copy()
    for j in range(1, m):
        last = last + Y[j]
    vec[N - 1] = jreshape(last, (dims[N - 1], out_rks[N - 1], out_rks[N]))
    return TT(vec, dims, out_rks, ot)


def ttrand_rounding_residual(H: TTO, y: TT, b: TT, rks, draws, orthogonal: bool = True,
                             block_rks: Optional[int] = None) -> TT:
    """Rounding of ``H*y - b`` (``src/tt_randtools.jl:169-214``)."""
    N = y.N
    dims = y.dims
    WAy, srk = tt_recursive_sketch_op(H, y, rks, draws, orthogonal, block_rks)
    Wb, srk2 = tt_recursive_sketch(b, rks, draws, orthogonal, block_rks)
    assert srk == srk2
    vec: List[Optional[np.ndarray]] = [None] * N
    out_rks = [1] * (N + 1)
    ot = [0] * N
    Ay = np.einsum("jab,ijcd->iacbd", y.cores[0], H.cores[0], optimize=True)[:, :, 0, :, :]  # (n,1,Ry2,RA2)
    bk = jreshape(b.cores[0], (dims[0], 1, b.rks[1]))
    for k in range(N - 1):
        Z = np.einsum("irab,abs->irs", Ay, WAy[k + 1], optimize=True) - np.tensordot(bk, Wb[k + 1], axes=([2], [0]))
        Z = jreshape(Z, (dims[k] * out_rks[k], srk[k + 1]))
        Q = _qr_thin(Z)
        out_rks[k + 1] = Q.shape[1]
        vec[k] = jreshape(Q, (dims[k], out_rks[k], out_rks[k + 1]))
        ot[k] = 1
        P = np.einsum("irab,irs->sab", Ay, vec[k], optimize=True)
        Ay = _op_project(P, y.cores[k + 1], H.cores[k + 1])
        Pb = np.einsum("ira,irs->sa", bk, vec[k], optimize=True)
        bk = np.einsum("sa,iab->isb", Pb, b.cores[k + 1], optimize=True)
    out_rks[N] = 1
    vec[N - 1] = jreshape(Ay, (dims[N - 1], out_rks[N - 1], out_rks[N])) - jreshape(bk, (dims[N - 1], out_rks[N - 1], out_rks[N]))
    return TT(vec, dims, out_rks, ot)


def rel_err(x: TT, y: TT) -> float:
    """``norm(x - y) / norm(x)`` through the QR-based TT norm (the error metric
    of ``examples/quantics_rounding_benchmark.jl:70-72``)."""
    nx = tt_norm(orthogonalize(x, i=x.N))
    return tt_norm(tt_sub(x, y)) / nx


# ---------------------------------------------------------------------------
# Sketched Krylov basis and sketched Rayleigh-Ritz: the solver-side callers of the
# boundary sketches and the randomized rounding (src/tt_solvers.jl:330-689)
# ---------------------------------------------------------------------------

def _eig_sorted(M: np.ndarray, B: Optional[np.ndarray] = None):
    """Julia's ``eigen(M)`` / ``eigen(M, B)``: eigenvalues sorted by (real, imag)
    (``LinearAlgebra.eigsortby``), eigenvectors in the same order."""
    if B is None:
        lam, V = np.linalg.eig(M)
    else:
        import scipy.linalg
        lam, V = scipy.linalg.eig(M, B)
    order = sorted(range(len(lam)), key=lambda i: (lam[i].real, lam[i].imag))
    return lam[order], V[:, order]


def expand_basis(H: TTO, B_window: List[TT], W_B_window: List[List[np.ndarray]], B_sketch_window: List[np.ndarray],
                 rmax: int, s: int, draws, draws_extra, orthogonal: bool = True, block_rks: Optional[int] = 8,
                 k_trunc: Optional[int] = None):
    """``expand_basis!`` (``src/tt_solvers.jl:354-445``), one sketch-orthogonal
    Arnoldi step.  The combined sketch of ``H b_prev`` (ranks ``2 rmax``, boundary
    sketch of size ``s``; ``draws`` / ``draws_extra`` are the draws of ``seed`` /
    ``seed + 1``) gives the least-squares coefficients ``h`` against the window's
    boundary sketches (``:376``); the left-to-right sweep over the ``1 + len(window)``
    terms ``H b_prev, -h_i b_i`` with the precomputed sketches ``W_HB`` and
    ``W_B_window`` (``:382-437``); ``tt_rounding(rmax)`` (``:439``); the combined
    sketch of the result, normalised by the norm of its boundary sketch
    (``:442-451``); the three windows mutated as a sliding window of ``k_trunc``
    (``:453-455``).  Returns ``(b_new, sketch_b_new, sketch_Hb_prev)``."""
    if k_trunc is None:
        k_trunc = len(B_window)
    b_prev = B_window[-1]
    dims, N = b_prev.dims, b_prev.N
    rks = [2 * rmax] * (N + 1)
    rks[N] = 1
    sketch_Hb_prev, WHb_raw, s_rks = tt_combined_sketch(b_prev, rks, s, draws, draws_extra, orthogonal, block_rks, H=H)
    W_HB = [jreshape(WHb_raw[k], (b_prev.rks[k] * H.rks[k], s_rks[k])) for k in range(N + 1)]
    # h = hcat(B_sketch_window...) \ sketch_Hb_prev (a tall least-squares solve)
    h = np.linalg.lstsq(np.column_stack(B_sketch_window), sketch_Hb_prev, rcond=None)[0]
    m = 1 + len(B_window)
    vec_out: List[Optional[np.ndarray]] = [None] * N
    out_rks = [1] * (N + 1)
    ot = [0] * N
    b1 = jreshape(b_prev.cores[0], (dims[0], b_prev.rks[1]))
    H1 = jreshape(H.cores[0], (dims[0], dims[0], H.rks[1]))
    Ay1 = np.einsum("ijb,ja->iab", H1, b1, optimize=True)  # (i1, alpha2, beta2)
    Y = [jreshape(Ay1, (dims[0], 1, b_prev.rks[1] * H.rks[1]))]
    for i, b in enumerate(B_window):
        Y.append(-h[i] * b.cores[0])
    for k in range(N - 1):
        Z = np.tensordot(Y[0], W_HB[k + 1], axes=([2], [0]))
        for i, W in enumerate(W_B_window):
            Z = Z + np.tensordot(Y[1 + i], W[k + 1], axes=([2], [0]))
        Q = _qr_thin(jreshape(Z, (dims[k] * out_rks[k], s_rks[k + 1])))
        out_rks[k + 1] = Q.shape[1]
        vec_out[k] = jreshape(Q, (dims[k], out_rks[k], out_rks[k + 1]))
        ot[k] = 1
        y1k = jreshape(Y[0], (dims[k], out_rks[k], b_prev.rks[k + 1], H.rks[k + 1]))
        P = np.einsum("irab,irs->sab", y1k, vec_out[k], optimize=True)
        t = _op_project(P, b_prev.cores[k + 1], H.cores[k + 1])
        Ynew = [jreshape(t, (dims[k + 1], out_rks[k + 1], b_prev.rks[k + 2] * H.rks[k + 2]))]
        for i, b in enumerate(B_window):
            Pi = np.einsum("ira,irs->sa", Y[1 + i], vec_out[k], optimize=True)
            Ynew.append(np.einsum("sa,iab->isb", Pi, b.cores[k + 1], optimize=True))
        Y = Ynew
    last = Y[0].copy()
    for j in range(1, m):
        last = last + Y[j]
    vec_out[N - 1] = jreshape(last, (dims[N - 1], out_rks[N - 1], 1))
    b_new = tt_rounding(TT(vec_out, dims, out_rks, ot), rmax=rmax)
    sketch_b_new, W_B_new, _ = tt_combined_sketch(b_new, rks, s, draws, draws_extra, orthogonal, block_rks)
    beta = float(np.linalg.norm(sketch_b_new))
    sketch_b_new = sketch_b_new / beta
    root = b_new.ot.index(0)  # W_B_new[i] ./= beta for i = 2:findfirst(ot == 0) (1-based)
    for i in range(1, root + 1):
        W_B_new[i] = W_B_new[i] / beta
    b_new = tt_scale(1.0 / beta, b_new)
    for win, item in ((B_window, b_new), (B_sketch_window, sketch_b_new), (W_B_window, W_B_new)):
        win.append(item)
        if len(win) > k_trunc:
            win.pop(0)
    return b_new, sketch_b_new, sketch_Hb_prev


def sketched_rr(H: TTO, B: Sequence[TT], C: np.ndarray, D: np.ndarray, rmax: int, draws_round, draws_sample,
                orthogonal: bool = True, block_rks: Optional[int] = 8, kappa_max: Optional[float] = None,
                stable: bool = False):
    """``sketched_rr`` (``src/tt_solvers.jl:466-514``), sketched Rayleigh-Ritz
    (Tropp & Nakatsukasa 2022) on the basis ``B`` with sketches ``C = Omega B``,
    ``D = Omega H B``: ``stable=false`` solves ``C M = D`` in the least-squares
    sense (column-pivoted QR) and takes ``eigen(M)``; ``stable=true`` whitens with
    the thin SVD of ``C``, drops directions with ``sigma/sigma_1 <= 1/kappa_max``
    and solves ``U^T D V_k z = lambda diag(sigma_k) z``.  The Ritz vector is the
    randomized rounding of ``sum_j y_j B_j`` to ranks ``4 rmax`` (``draws_round``),
    SVD-rounded at ``tol = res_sketch^2/2`` and normalised; the sampled residual
    compares the boundary sketches of ``H X`` and ``X`` of size ``2N`` drawn from
    ``draws_sample``.  Returns ``(lambdas, X, res_sketch, res_sample)``."""
    N = B[0].N
    if kappa_max is None:
        kappa_max = 1.0 / math.sqrt(np.finfo(np.float64).eps)
    if stable:
        U, sig, Vt = np.linalg.svd(C, full_matrices=False)
        keep = [i for i in range(len(sig)) if sig[i] / sig[0] > 1.0 / kappa_max]
        k = keep[-1] + 1 if keep else 1
        Uk, Vk, sk = U[:, :k], Vt[:k].T, sig[:k]
        M = Uk.T @ D @ Vk
        lam, Z = _eig_sorted(M, np.diag(sk))
        Yv = np.real(Vk @ Z)
        for j in range(k):
            Yv[:, j] = Yv[:, j] / np.linalg.norm(C @ Yv[:, j])
    else:
        M = np.linalg.lstsq(C, D, rcond=None)[0]
        lam, Yv = _eig_sorted(M)
        Yv = np.real(Yv)
    lam = np.real(lam)
    res_sketch = float(np.linalg.norm(D @ Yv[:, 0] - lam[0] * (C @ Yv[:, 0])))
    y = Yv[:, 0] / np.linalg.norm(Yv[:, 0])
    rks = [1] + [4 * rmax] * (N - 1) + [1]
    X = ttrand_rounding_lincomb(list(y), list(B), rks, draws_round, orthogonal, block_rks)
    X = tt_rounding(X, tol=res_sketch ** 2 / 2)
    # X / norm(X): when res_sketch^2/2 exceeds norm(X) the rounding returned
    # zeros_tt (src/tt_rounding.jl:176-178) and the reference's 1/0 = Inf turns
    # the Ritz vector into NaNs; restated as such
    with np.errstate(divide="ignore", invalid="ignore"):
        X = tt_scale(np.float64(1.0) / np.float64(tt_norm(X)), X)
        S1, _ = tt_sketch(X, 2 * N, draws_sample, orthogonal, block_rks, H=H)
        S2, _ = tt_sketch(X, 2 * N, draws_sample, orthogonal, block_rks)
        res_sample = float(np.linalg.norm(S1 - lam[0] * S2) / np.linalg.norm(S2))
    return lam, X, res_sketch, res_sample


def sketched_rayleigh_ritz(H: TTO, b0: TT, d: int, rmax: int, draws_for, orthogonal: bool = True,
                           block_rks: Optional[int] = 8, seed: int = 1234, sketch_size: Optional[int] = None,
                           k_trunc: Optional[int] = None, e_nuc: float = 0.0, sample_seeds=None):
    """``sketched_rayleigh_ritz(H, b0, d, rmax)`` (``src/tt_solvers.jl:548-620``):
    a ``d``-dimensional sketch-orthogonal Krylov basis from ``b0`` (normalised by
    its boundary sketch), with a sketched Rayleigh-Ritz after every expansion and
    a final one over all ``d`` vectors.  ``draws_for(seed)`` returns the draws of
    every sketch made with that seed (per core, long enough for any block
    structure; ``seed + 1`` feeds the combined sketches' extensions);
    ``sample_seeds`` are the seeds the reference draws with ``rand(Int)`` for the
    sampled residuals.  Returns ``(e_sketch, psi,
    history)``; history entries are dicts with ``iter, e_sketch, e_true,
    res_sketch, res_sample``."""
    N = b0.N
    if sketch_size is None:
        sketch_size = 4 * d
    if k_trunc is None:
        k_trunc = d - 1
    rks = [2 * rmax] * (N + 1)
    rks[N] = 1
    samp = list(sample_seeds) if sample_seeds is not None else [None] * d

    def comb(x, Hx, s):
        return tt_combined_sketch(x, rks, s, draws_for(seed), draws_for(seed + 1), orthogonal, block_rks, H=Hx)

    sk1, W_B_1, _ = comb(b0, None, sketch_size)
    s = len(sk1)
    beta = float(np.linalg.norm(sk1))
    sk1 = sk1 / beta
    root = b0.ot.index(0)
    for i in range(1, root + 1):
        W_B_1[i] = W_B_1[i] / beta
    B = [tt_scale(1.0 / beta, b0)]
    sketch_b, sketch_Hb = [sk1], []
    B_window, W_B_window, B_sketch_window = [B[0]], [W_B_1], [sk1]
    history = []

    def srr(Bs, sb, sHb, ss):
        C = np.column_stack(sb)
        D = np.column_stack(sHb)
        return sketched_rr(H, Bs, C, D, rmax, draws_for(seed), draws_for(ss), orthogonal, block_rks)

    for j in range(2, d + 1):
        b_new, sk_new, sk_Hb = expand_basis(H, B_window, W_B_window, B_sketch_window, rmax, s,
                                            draws_for(seed), draws_for(seed + 1), orthogonal, block_rks, k_trunc)
        B.append(b_new)
        sketch_b.append(sk_new)
        sketch_Hb.append(sk_Hb)
        lam, psi, r1, r2 = srr(B[:j - 1], sketch_b[:j - 1], sketch_Hb[:j - 1], samp[j - 2])
        with np.errstate(invalid="ignore"):
            e_true = dot_operator(psi, H, psi) + e_nuc
        history.append(dict(iter=j - 1, e_sketch=lam[0] + e_nuc, e_true=e_true, res_sketch=r1, res_sample=r2))
    skH, _, _ = comb(B[-1], H, s)
    sketch_Hb.append(skH)
    lam, psi, r1, r2 = srr(B, sketch_b, sketch_Hb, samp[d - 1])
    e = lam[0] + e_nuc
    with np.errstate(invalid="ignore"):
        e_true = dot_operator(psi, H, psi) + e_nuc
    history.append(dict(iter=d, e_sketch=e, e_true=e_true, res_sketch=r1, res_sample=r2))
    return e, psi, history


# ---------------------------------------------------------------------------
# 1D Hubbard chain MPO with local dimension 4 (site = (up, down) pair)
# ---------------------------------------------------------------------------

def _site_ops():
    # local basis index = n_up + 2*n_dn  (0: empty, 1: up, 2: down, 3: both)
    cu = np.zeros((4, 4))
    cd = np.zeros((4, 4))
    # c_up |n_up n_dn>: removes the up electron; the up mode is ordered first
    cu[0, 1] = 1.0
    cu[2, 3] = 1.0
    # c_dn acts after passing the up mode: sign (-1)^{n_up}
    cd[0, 2] = 1.0
    cd[1, 3] = -1.0
    F = np.diag([1.0, -1.0, -1.0, 1.0])  # (-1)^{n_up + n_dn}
    nud = np.diag([0.0, 0.0, 0.0, 1.0])
    I = np.eye(4)
    return I, cu, cd, F, nud


def hubbard_1d_mpo(L: int, t: float = 1.0, U: float = 4.0) -> TTO:
    """Closed-form MPO (bond dimension 6) of the open 1D Hubbard chain
    ``H = -t sum_{i,s} (c+_{i,s} c_{i+1,s} + h.c.) + U sum_i n_{i,up} n_{i,dn}``
    with Jordan-Wigner strings, one site per core with local dimension 4.

    Convention note: the reference builds this Hamiltonian from
    ``hubbard_1D`` + ``hV_to_mpo`` (``src/models.jl:403-419,180-231``) on 2L
    spin-orbital sites of dimension 2, whose interaction term carries the
    factor ``V[i]*(term + h.c.)``, i.e. ``2U n_up n_dn``.  This builder uses the
    textbook ``U n_up n_dn``; pass ``U=2U`` to reproduce the reference's
    Hamiltonian.  Checked against a dense Hamiltonian in the tests."""
    I, cu, cd, F, nud = _site_ops()
    D = 6
    # W[a, b] operator-valued; H = W_1 W_2 ... W_L with boundary row 5 / column 0
    # hop c+_{i,s} c_{i+1,s} = [c+_s F]_i [c_s]_{i+1}; its h.c. = [F c_s]_i [c+_s]_{i+1}
    W = np.zeros((D, D, 4, 4))
    W[0, 0] = I
    W[1, 0] = cu
    W[2, 0] = cu.T
    W[3, 0] = cd
    W[4, 0] = cd.T
    W[5, 0] = U * nud
    W[5, 1] = -t * (cu.T @ F)
    W[5, 2] = -t * (F @ cu)
    W[5, 3] = -t * (cd.T @ F)
    W[5, 4] = -t * (F @ cd)
    W[5, 5] = I
    cores = []
    for i in range(L):
        if i == 0:
            Wi = W[5:6, :, :, :]
        elif i == L - 1:
            Wi = W[:, 0:1, :, :]
        else:
            Wi = W
        # TT operator core [i_out, j_in, a, b]
        cores.append(np.transpose(Wi, (2, 3, 0, 1)).copy())
    rks = [1] + [D] * (L - 1) + [1]
    return TTO(cores, tuple([4] * L), rks, [0] * L)


def hubbard_1d_dense(L: int, t: float = 1.0, U: float = 4.0) -> np.ndarray:
    """Dense ``4^L x 4^L`` Hubbard Hamiltonian from explicit fermion operators
    on 2L modes (mode order up_1, dn_1, up_2, ...), in the column-major
    multi-index convention of ``tto_to_tensor`` (site 1 fastest)."""
    nm = 2 * L
    dim = 2 ** nm

    def ann(mode):
        A = np.zeros((dim, dim))
        for s in range(dim):
            if (s >> mode) & 1:
                sign = (-1) ** bin(s & ((1 << mode) - 1)).count("1")
                A[s ^ (1 << mode), s] = sign
        return A

    c = [ann(m) for m in range(nm)]
    H = np.zeros((dim, dim))
    for i in range(L - 1):
        for sp in range(2):
            a, b = 2 * i + sp, 2 * (i + 1) + sp
            hop = c[a].T @ c[b]
            H += -t * (hop + hop.T)
    for i in range(L):
        H += U * (c[2 * i].T @ c[2 * i]) @ (c[2 * i + 1].T @ c[2 * i + 1])
    # bit layout: bit 2i = up on site i, bit 2i+1 = dn; local index n_up + 2 n_dn
    # => state integer s is exactly sum_i local_i * 4^i (site 1 fastest)
    return H
