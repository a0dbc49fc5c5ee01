"""Host mirror of the TensorTrains rounding API over libttround.

Names, argument meanings, defaults and error behaviour follow the Julia
package (``src/tt_rounding.jl``, ``src/tt_randtools.jl``, ``src/sketches.jl``,
``src/tt_operations.jl``, ``src/tt_tools.jl``); every numerical operation runs
in the HIP library.  Two object kinds exist:

* ``TTvector`` / ``TToperator``: host objects with the reference's fields
  (``ttv_vec`` cores of shape ``(n_k, r_{k-1}, r_k)``, ``ttv_dims``,
  ``ttv_rks``, ``ttv_ot``); each call uploads, computes on the GPU and
  downloads the result, like the reference's value semantics.
* ``DeviceTT``: a TT resident in device memory, for batched pipelines that keep
  data in HBM between calls (``round_rand_batch``, ``round_svd_batch``).
"""
from __future__ import annotations

import ctypes as C
import math
from typing import List, Optional, Sequence

import numpy as np

from . import _lib as L


# --------------------------------------------------------------------- helpers

def _wrapped_prod(vals) -> int:
    p = 1
    for v in vals:
        p = (p * int(v)) & 0xFFFFFFFFFFFFFFFF
    return p - (1 << 64) if p >= (1 << 63) else p


def r_and_d_to_rks(rks, dims, rmax: int = 1024) -> List[int]:
    """``r_and_d_to_rks`` (``src/tt_rounding.jl:6-24``), integer bookkeeping on the host."""
    d = len(dims)
    out = [1] * len(rks)
    for i in range(d):
        right, left = _wrapped_prod(dims[i:]), _wrapped_prod(dims[:i])
        v = min(int(rks[i]), int(rmax))
        if right > 0:
            v = min(v, right)
        if left > 0:
            v = min(v, left)
        out[i] = v
    return out


def _block_ptrs(blocks, keep):
    """Pointer array over per-core FP64 block arrays (column-major copies kept alive in ``keep``)."""
    if blocks is None:
        return C.cast(None, L.f64pp)
    arrs = [np.asfortranarray(np.asarray(b, dtype=np.float64)) for b in blocks]
    keep.extend(arrs)
    ptrs = (L.f64p * len(arrs))(*[a.ctypes.data_as(L.f64p) for a in arrs])
    keep.append(ptrs)
    return C.cast(ptrs, L.f64pp)


def _opts(orthogonal=True, sketch_prec="f64", block_rks=None, seed=1234, p=None, oversampling=1.0, forward=False,
          blocks=None, blocks_right=None, sketch_accum="sketch"):
    """``blocks``: optional per-core sketch blocks exactly as ``generate_sketch_blocks``
    returns them (``src/sketches.jl:44-84``), used instead of the library's draws.
    ``sketch_accum``: ``"sketch"`` runs the sketch contractions in ``sketch_prec``
    (FP32 accumulation for FP32/BF16); ``"f64"`` rounds only the blocks to
    ``sketch_prec`` and contracts in FP64."""
    o = L.RandOpts()
    o.forward = 1 if forward else 0
    o.orthogonal = 1 if orthogonal else 0
    o.sketch_prec = L.PREC[sketch_prec] if isinstance(sketch_prec, str) else int(sketch_prec)
    o.block_rks = int(block_rks) if block_rks else 0
    o.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    keep = []
    if p is not None:
        pa = L.as_i64(p)
        keep.append(pa)
        o.p = L.ptr_i64(pa)
    else:
        o.p = C.cast(None, L.i64p)
    o.oversampling = float(oversampling)
    o.blocks = _block_ptrs(blocks, keep)
    o.blocks_right = _block_ptrs(blocks_right, keep)
    if sketch_accum not in ("sketch", "f64"):
        raise ValueError('sketch_accum must be "sketch" or "f64"')
    o.sketch_accum = 1 if sketch_accum == "f64" else 0
    return o, keep


# --------------------------------------------------------------------- device objects

class DeviceTT:
    """A TT held in device memory by a libttround handle."""

    def __init__(self, handle: L.Handle, tid: int):
        self.handle = handle
        self.id = int(tid)
        d = C.c_int64()
        handle.check(handle.lib.ttr_tt_info(handle.h, self.id, C.byref(d), None, None, None))
        self.d = d.value
        dims = np.zeros(self.d, np.int64)
        rks = np.zeros(self.d + 1, np.int64)
        ot = np.zeros(self.d, np.int64)
        handle.check(handle.lib.ttr_tt_info(handle.h, self.id, None, L.ptr_i64(dims), L.ptr_i64(rks), L.ptr_i64(ot)))
        self.dims = tuple(int(x) for x in dims)
        self.rks = [int(x) for x in rks]
        self.ot = [int(x) for x in ot]

    # construction
    @classmethod
    def upload(cls, x: "TTvector", handle: Optional[L.Handle] = None) -> "DeviceTT":
        h = handle or L.default_handle()
        d = x.N
        cores = [np.asfortranarray(np.asarray(c, dtype=np.float64)) for c in x.ttv_vec]
        for k, c in enumerate(cores):
            exp = (x.ttv_dims[k], x.ttv_rks[k], x.ttv_rks[k + 1])
            if c.shape != exp:
                raise L.DimensionMismatch(f"core {k} has shape {c.shape}, expected {exp}")
        arr = (L.f64p * d)(*[L.ptr_f64(c) for c in cores])
        dims, rks, ot = L.as_i64(x.ttv_dims), L.as_i64(x.ttv_rks), L.as_i64(x.ttv_ot)
        tid = C.c_int64()
        h.check(h.lib.ttr_tt_upload(h.h, d, L.ptr_i64(dims), L.ptr_i64(rks), L.ptr_i64(ot), arr, C.byref(tid)))
        return cls(h, tid.value)

    @classmethod
    def randn(cls, dims, rks, seed: int, handle: Optional[L.Handle] = None) -> "DeviceTT":
        """Seeded N(0,1) TT drawn on the device (``rand_tt`` semantics, ``src/tt_tools.jl:255-303``)."""
        h = handle or L.default_handle()
        d = len(dims)
        if isinstance(rks, (int, np.integer)):
            rks = [int(rks)] * (d + 1)
        da, ra = L.as_i64(dims), L.as_i64(rks)
        tid = C.c_int64()
        h.check(h.lib.ttr_tt_randn(h.h, d, L.ptr_i64(da), L.ptr_i64(ra), int(seed) & 0xFFFFFFFFFFFFFFFF, C.byref(tid)))
        return cls(h, tid.value)

    def core(self, k: int) -> np.ndarray:
        shape = (self.dims[k], self.rks[k], self.rks[k + 1])
        out = np.empty(shape, dtype=np.float64, order="F")
        self.handle.check(self.handle.lib.ttr_tt_download_core(self.handle.h, self.id, int(k), L.ptr_f64(out)))
        return out

    def to_host(self) -> "TTvector":
        return TTvector(self.d, [self.core(k) for k in range(self.d)], self.dims, list(self.rks), list(self.ot))

    def free(self):
        if self.id is not None and self.handle.h:
            self.handle.check(self.handle.lib.ttr_tt_free(self.handle.h, self.id))
        self.id = None

    def norm(self) -> float:
        out = np.zeros(1)
        ids = L.as_i64([self.id])
        self.handle.check(self.handle.lib.ttr_norm(self.handle.h, 1, L.ptr_i64(ids), L.ptr_f64(out)))
        return float(out[0])


class DeviceTTO:
    """A TT operator held in device memory."""

    def __init__(self, A: "TToperator", handle: Optional[L.Handle] = None):
        h = handle or L.default_handle()
        self.handle = h
        cores = [np.asfortranarray(np.asarray(c, dtype=np.float64)) for c in A.tto_vec]
        arr = (L.f64p * A.N)(*[L.ptr_f64(c) for c in cores])
        dims, rks = L.as_i64(A.tto_dims), L.as_i64(A.tto_rks)
        tid = C.c_int64()
        h.check(h.lib.ttr_tto_upload(h.h, A.N, L.ptr_i64(dims), L.ptr_i64(rks), arr, C.byref(tid)))
        self.id = tid.value
        self.dims = tuple(A.tto_dims)
        self.rks = list(A.tto_rks)

    def free(self):
        if self.id is not None and self.handle.h:
            self.handle.check(self.handle.lib.ttr_tt_free(self.handle.h, self.id))
        self.id = None


def lincomb_device(coeffs: Sequence[float], xs: Sequence[DeviceTT]) -> DeviceTT:
    h = xs[0].handle
    a = np.ascontiguousarray(np.asarray(coeffs, dtype=np.float64))
    ids = L.as_i64([x.id for x in xs])
    out = C.c_int64()
    h.check(h.lib.ttr_lincomb(h.h, len(xs), L.ptr_f64(a), L.ptr_i64(ids), C.byref(out)))
    return DeviceTT(h, out.value)


def mpo_apply_device(A: DeviceTTO, x: DeviceTT) -> DeviceTT:
    h = x.handle
    out = C.c_int64()
    h.check(h.lib.ttr_mpo_apply(h.h, A.id, x.id, C.byref(out)))
    return DeviceTT(h, out.value)


def hadamard_device(xs: Sequence[DeviceTT], first_fastest: bool = False) -> DeviceTT:
    """Entrywise product of device TTs (``ttr_hadamard``): per mode index the
    Kronecker product of the core slices.  ``first_fastest=False`` is the
    reference's ``x * y`` (``src/tt_operations.jl:273-290``, the last factor's
    rank index fastest); ``True`` is the layout of the NTuple methods."""
    h = xs[0].handle
    ids = L.as_i64([x.id for x in xs])
    out = C.c_int64()
    h.check(h.lib.ttr_hadamard(h.h, len(xs), L.ptr_i64(ids), 1 if first_fastest else 0, C.byref(out)))
    return DeviceTT(h, out.value)


def _hadamard_dev(ys) -> DeviceTT:
    """The implicit product of an NTuple argument, built on the device with the
    first factor's rank index fastest (the reference's order for these methods)."""
    ys = list(ys)
    if not ys or any(not isinstance(t, TTvector) for t in ys):
        raise TypeError("a tuple argument must hold TTvectors")
    if any(t.ttv_dims != ys[0].ttv_dims for t in ys):
        raise L.DimensionMismatch("all TTs of the product must share dims")
    devs = [t._dev() for t in ys]
    if len(devs) == 1:
        return devs[0]
    try:
        return hadamard_device(devs, first_fastest=True)
    finally:
        for d in devs:
            d.free()


class _Product:
    """An NTuple argument of the sketch functions: the entrywise product of the
    factors, built on the device at each use (``_dev``) and never downloaded."""

    def __init__(self, ys):
        self.factors = tuple(ys)
        self.N = self.factors[0].N
        self.ttv_dims = self.factors[0].ttv_dims
        self.factor_rks = [[t.ttv_rks[k] for t in self.factors] for k in range(self.N + 1)]
        self.ttv_rks = [int(np.prod(r)) for r in self.factor_rks]

    def _dev(self) -> DeviceTT:
        return _hadamard_dev(self.factors)

    def split(self, W, k):
        """Sketch rows of bond ``k`` as the reference's ``(r_1, ..., r_M, s)`` array."""
        return W.reshape(tuple(self.factor_rks[k]) + (W.shape[-1],), order="F")


def orthogonalize_batch(xs: Sequence[DeviceTT], i: int = 1) -> List[DeviceTT]:
    h = xs[0].handle
    ids = L.as_i64([x.id for x in xs])
    out = np.zeros(len(xs), np.int64)
    h.check(h.lib.ttr_orthogonalize(h.h, len(xs), L.ptr_i64(ids), int(i), L.ptr_i64(out)))
    return [DeviceTT(h, t) for t in out]


def round_svd_batch(xs: Sequence[DeviceTT], tol: float = 1e-12, rmax: int = 2 ** 14,
                    direction: str = "left") -> List[DeviceTT]:
    h = xs[0].handle
    if direction not in L.DIRECTION:
        raise ValueError("Invalid direction. Use direction=:right or direction=:left")
    ids = L.as_i64([x.id for x in xs])
    out = np.zeros(len(xs), np.int64)
    h.check(h.lib.ttr_round_svd(h.h, len(xs), L.ptr_i64(ids), float(tol), int(rmax), L.DIRECTION[direction],
                                L.ptr_i64(out)))
    return [DeviceTT(h, t) for t in out]


def norm_batch(xs: Sequence[DeviceTT]) -> np.ndarray:
    h = xs[0].handle
    ids = L.as_i64([x.id for x in xs])
    out = np.zeros(len(xs))
    h.check(h.lib.ttr_norm(h.h, len(xs), L.ptr_i64(ids), L.ptr_f64(out)))
    return out


def dot_batch(xs: Sequence[DeviceTT], ys: Sequence[DeviceTT]) -> np.ndarray:
    """``<xs[i], ys[i]>`` for every pair, all pairs in the same launches
    (``dot``, ``src/tt_operations.jl:302-321``)."""
    if len(xs) != len(ys) or not xs:
        raise ValueError("xs and ys must be non-empty and of equal length")
    h = xs[0].handle
    a, b = L.as_i64([x.id for x in xs]), L.as_i64([y.id for y in ys])
    out = np.zeros(len(xs))
    h.check(h.lib.ttr_dot(h.h, len(xs), L.ptr_i64(a), L.ptr_i64(b), L.ptr_f64(out)))
    return out


def dot_operator_batch(psis: Sequence[DeviceTT], ops, phis: Sequence[DeviceTT]) -> np.ndarray:
    """``<psis[i]|ops[i]|phis[i]>`` for every triple (``dot_operator``,
    ``src/tt_operations.jl:385-421``); ``ops`` is one DeviceTTO or one per triple."""
    if isinstance(ops, DeviceTTO):
        ops = [ops] * len(psis)
    if not (len(psis) == len(ops) == len(phis)) or not psis:
        raise ValueError("psis, ops and phis must be non-empty and of equal length")
    h = psis[0].handle
    a, o, b = L.as_i64([x.id for x in psis]), L.as_i64([A.id for A in ops]), L.as_i64([y.id for y in phis])
    out = np.zeros(len(psis))
    h.check(h.lib.ttr_dot_operator(h.h, len(psis), L.ptr_i64(a), L.ptr_i64(o), L.ptr_i64(b), L.ptr_f64(out)))
    return out


def round_rand_batch(items, orthogonal: bool = True, seed: int = 1234, block_rks: Optional[int] = None,
                     sketch_prec: str = "f64", p=None, oversampling: float = 1.0, blocks=None,
                     sketch_accum: str = "sketch") -> List[DeviceTT]:
    """Batched randomized rounding of linear combinations.

    ``items`` is a list of ``(alphas, [DeviceTT...], rks)`` triples, or of
    ``(alphas, DeviceTTO, [DeviceTT...], rks)`` quadruples for
    ``alpha_1 H y_1 + sum_{j>1} alpha_j y_j`` (implicit operator product)."""
    h = items[0][-2][0].handle
    keep = []
    arr = (L.RandItem * len(items))()
    for i, item in enumerate(items):
        if len(item) == 4:
            alpha, op, terms, rks = item
        else:
            (alpha, terms, rks), op = item, None
        a = np.ascontiguousarray(np.asarray(alpha, dtype=np.float64))
        ids = L.as_i64([t.id for t in terms])
        r = L.as_i64(rks)
        keep += [a, ids, r]
        arr[i].nterms = len(terms)
        arr[i].alpha = L.ptr_f64(a)
        arr[i].term_ids = L.ptr_i64(ids)
        arr[i].op_id = -1 if op is None else op.id
        arr[i].rks = L.ptr_i64(r)
    o, k2 = _opts(orthogonal, sketch_prec, block_rks, seed, p, oversampling, blocks=blocks, sketch_accum=sketch_accum)
    out = np.zeros(len(items), np.int64)
    h.check(h.lib.ttr_round_rand(h.h, len(items), arr, C.byref(o), L.ptr_i64(out)))
    return [DeviceTT(h, t) for t in out]


def round_rand_tuple_batch(items, orthogonal: bool = True, seed: int = 1234, block_rks: Optional[int] = None,
                           sketch_prec: str = "f64", p=None, oversampling: float = 1.0, blocks=None,
                           sketch_accum: str = "sketch") -> List[DeviceTT]:
    """Batched randomized rounding of entrywise products, never formed
    (``ttr_round_rand_tuple``).  ``items`` is a list of ``([DeviceTT...], rks)``
    pairs: the factors of one product and its target sketch ranks."""
    h = items[0][0][0].handle
    nf = L.as_i64([len(f) for f, _ in items])
    ids = L.as_i64([t.id for f, _ in items for t in f])
    rk = L.as_i64([int(v) for _, r in items for v in r])
    o, keep = _opts(orthogonal, sketch_prec, block_rks, seed, p, oversampling, blocks=blocks,
                    sketch_accum=sketch_accum)
    out = np.zeros(len(items), np.int64)
    h.check(h.lib.ttr_round_rand_tuple(h.h, len(items), L.ptr_i64(nf), L.ptr_i64(ids), L.ptr_i64(rk), C.byref(o),
                                       L.ptr_i64(out)))
    return [DeviceTT(h, t) for t in out]


def _tuple_sketch(factors, rks, orthogonal, seed, block_rks, p, oversampling, sketch_prec, blocks=None,
                  forward=False):
    """Recursive sketch of the entrywise product of ``factors`` (TTvectors), never
    formed (``ttr_sketch_tuple``), right to left or (``forward``, FP64) left to
    right: ``W[k]`` of shape ``(r_1k, ..., r_Mk, s_k)``."""
    factors = list(factors)
    N = factors[0].N
    if any(t.ttv_dims != factors[0].ttv_dims for t in factors):
        raise L.DimensionMismatch("all TTs of the product must share dims")
    devs = [t._dev() for t in factors]
    h = devs[0].handle
    try:
        o, keep = _opts(orthogonal, sketch_prec, block_rks, seed, p, oversampling, blocks=blocks, forward=forward)
        ra = L.as_i64(rks)
        structure = block_structure_forward if forward else block_structure
        bv, pp = structure(factors[0].ttv_dims, rks, block_rks, p, oversampling)
        srk = [bv[k] * pp[k] for k in range(N + 1)]
        fr = [[t.ttv_rks[k] for t in factors] for k in range(N + 1)]
        Ws = [np.zeros((int(np.prod(fr[k])), srk[k]), order="F") for k in range(N + 1)]
        arr = (L.f64p * (N + 1))(*[L.ptr_f64(w) for w in Ws])
        out_srk = np.zeros(N + 1, np.int64)
        ids = L.as_i64([d.id for d in devs])
        h.check(h.lib.ttr_sketch_tuple(h.h, len(devs), L.ptr_i64(ids), L.ptr_i64(ra), C.byref(o), arr,
                                       L.ptr_i64(out_srk)))
    finally:
        for d in devs:
            d.free()
    assert [int(v) for v in out_srk] == srk
    return [w.reshape(tuple(fr[k]) + (srk[k],), order="F") for k, w in enumerate(Ws)], srk


def block_structure(dims, rks, block_rks: Optional[int] = None, p=None, oversampling: float = 1.0):
    """Block ranks ``bv`` and block counts ``p`` of the reverse sketch
    (``src/sketches.jl:107-123,186-198``)."""
    d = len(dims)
    bv = [1] * (d + 1)
    for k in range(d):
        bv[k] = int(block_rks) if block_rks else d
    for k in range(d - 1, -1, -1):
        bv[k] = min(bv[k], int(dims[k]) * bv[k + 1])
    if p is None:
        pp = [int(math.ceil(int(r) / b)) for r, b in zip(rks, bv)]
        for k in range(1, d):
            pp[k] = max(pp[k], pp[k - 1])
        pp[d] = 1
    else:
        pp = [int(v) for v in p]
    if oversampling != 1:
        for k in range(d):
            pp[k] = int(math.ceil(oversampling * pp[k]))
    return bv, pp


def tt_sketch_batch(xs: Sequence[DeviceTT], s: int, ops=None, orthogonal: bool = True, seed: int = 1234,
                    block_rks: Optional[int] = None, sketch_prec: str = "f64", blocks=None):
    """Boundary-only reverse sketches of a batch (``tt_sketch``, ``src/sketches.jl:645-813``)
    in one library call: item ``i`` is ``xs[i]``, or ``ops[i] * xs[i]`` (never formed)
    when ``ops[i]`` is a ``DeviceTTO``.  All items share the draws of ``seed``.
    Returns ``(S, s_actual)`` with ``S[:, i]`` the sketch of item ``i``."""
    h = xs[0].handle
    ids = L.as_i64([x.id for x in xs])
    oids = L.as_i64([-1 if ops is None or ops[i] is None else ops[i].id for i in range(len(xs))])
    o, keep = _opts(orthogonal, sketch_prec, block_rks, seed, blocks=blocks)
    sa = np.zeros(1, np.int64)
    h.check(h.lib.ttr_tt_sketch(h.h, len(xs), L.ptr_i64(ids), L.ptr_i64(oids), int(s), C.byref(o),
                                C.cast(None, L.f64p), 0, L.ptr_i64(sa)))
    out = np.zeros((int(sa[0]), len(xs)), order="F")
    h.check(h.lib.ttr_tt_sketch(h.h, len(xs), L.ptr_i64(ids), L.ptr_i64(oids), int(s), C.byref(o),
                                L.ptr_f64(out), out.size, L.ptr_i64(sa)))
    return out, int(sa[0])


def block_structure_forward(dims, rks, block_rks: Optional[int] = None, p=None, oversampling: float = 1.0):
    """Block ranks and counts of the forward sketch (``src/sketches.jl:113-121,244-255``)."""
    d = len(dims)
    bv = [1] * (d + 1)
    for k in range(1, d + 1):
        bv[k] = int(block_rks) if block_rks else d
    for k in range(d):
        bv[k + 1] = min(bv[k + 1], int(dims[k]) * bv[k])
    if p is None:
        pp = [int(math.ceil(int(r) / b)) for r, b in zip(rks, bv)]
        for k in range(d - 1, 0, -1):
            pp[k] = max(pp[k], pp[k + 1])
        pp[0] = 1
    else:
        pp = [int(v) for v in p]
    if oversampling != 1:
        for k in range(1, d + 1):
            pp[k] = int(math.ceil(oversampling * pp[k]))
    return bv, pp


def sketch_blocks(dims, rks, k: int, orthogonal: bool = True, seed: int = 1234, block_rks: Optional[int] = None,
                  sketch_prec: str = "f64", p=None, oversampling: float = 1.0, forward: bool = False,
                  handle: Optional[L.Handle] = None) -> np.ndarray:
    """The sketch blocks of core ``k`` (0-based) as the device uses them, as the
    ``(bv_{k+1}, n_k, bv_k, p_k)`` array of ``generate_sketch_blocks`` (forward:
    ``(bv_k, n_k, bv_{k+1}, p_{k+1})``)."""
    h = handle or L.default_handle()
    d = len(dims)
    o, keep = _opts(orthogonal, sketch_prec, block_rks, seed, p, oversampling, forward)
    if forward:
        bv, pp = block_structure_forward(dims, rks, block_rks, p, oversampling)
        buf = np.zeros(int(bv[k] * dims[k] * bv[k + 1] * pp[k + 1]))
    else:
        bv, pp = block_structure(dims, rks, block_rks, p, oversampling)
        buf = np.zeros(int(bv[k + 1] * dims[k] * bv[k] * pp[k]))
    shape = np.zeros(4, np.int64)
    da, ra = L.as_i64(dims), L.as_i64(rks)
    h.check(h.lib.ttr_sketch_blocks(h.h, d, L.ptr_i64(da), L.ptr_i64(ra), C.byref(o), int(k), L.ptr_f64(buf),
                                    L.ptr_i64(shape)))
    return np.reshape(buf, tuple(int(s) for s in shape), order="F").copy()


def sketch_draws(k: int, count: int, seed: int = 1234, handle: Optional[L.Handle] = None) -> np.ndarray:
    h = handle or L.default_handle()
    out = np.zeros(int(count))
    h.check(h.lib.ttr_sketch_draws(h.h, int(k), int(count), int(seed) & 0xFFFFFFFFFFFFFFFF, L.ptr_f64(out)))
    return out


# --------------------------------------------------------------------- host objects

class TTvector:
    """``TTvector{Float64,N}`` (``src/tt_tools.jl:24-34``)."""

    def __init__(self, N, ttv_vec, ttv_dims, ttv_rks, ttv_ot=None):
        self.N = int(N)
        self.ttv_vec = [np.asarray(c, dtype=np.float64) for c in ttv_vec]
        self.ttv_dims = tuple(int(x) for x in ttv_dims)
        self.ttv_rks = [int(x) for x in ttv_rks]
        self.ttv_ot = [int(x) for x in (ttv_ot if ttv_ot is not None else [0] * self.N)]

    def copy(self) -> "TTvector":
        return TTvector(self.N, [c.copy() for c in self.ttv_vec], self.ttv_dims, list(self.ttv_rks), list(self.ttv_ot))

    def _dev(self) -> DeviceTT:
        return DeviceTT.upload(self)

    def __add__(self, other: "TTvector") -> "TTvector":
        return _host_lincomb([1.0, 1.0], [self, other], plain_sum=True)

    def __sub__(self, other: "TTvector") -> "TTvector":
        """``A - B = (-1.0)*B + A`` (``src/tt_operations.jl:453-455``)."""
        return tt_scale(-1.0, other) + self

    def __rmul__(self, a) -> "TTvector":
        return tt_scale(a, self)

    def __mul__(self, other):
        """``x * y`` (``src/tt_operations.jl:273-290``): the entrywise product,
        core slices ``kron(x_k[i], y_k[i])``; ``x * a`` for a number scales."""
        if not isinstance(other, TTvector):
            return tt_scale(other, self)
        if other.ttv_dims != self.ttv_dims:
            raise L.DimensionMismatch("Incompatible dimensions")
        xs = [self._dev(), other._dev()]
        out = hadamard_device(xs)
        res = out.to_host()
        for d in xs + [out]:
            d.free()
        return res

    def __truediv__(self, a) -> "TTvector":
        return tt_scale(1.0 / a, self)


class TToperator:
    """``TToperator{Float64,N}`` (``src/tt_tools.jl:61-71``)."""

    def __init__(self, N, tto_vec, tto_dims, tto_rks, tto_ot=None):
        self.N = int(N)
        self.tto_vec = [np.asarray(c, dtype=np.float64) for c in tto_vec]
        self.tto_dims = tuple(int(x) for x in tto_dims)
        self.tto_rks = [int(x) for x in tto_rks]
        self.tto_ot = [int(x) for x in (tto_ot if tto_ot is not None else [0] * self.N)]

    def __mul__(self, x: TTvector) -> TTvector:
        """``A*x`` (``src/tt_operations.jl:170-194``)."""
        op, xd = DeviceTTO(self), x._dev()
        dev = mpo_apply_device(op, xd)
        out = dev.to_host()
        for d in (dev, xd, op):
            d.free()
        return out


# --------------------------------------------------------------------- JSON I/O

def _json_obj(x):
    if isinstance(x, (str, bytes, bytearray)):
        import json
        return json.loads(x)
    return x


def _json_cores(flat, shapes, what):
    cores = []
    for k, (c, shp) in enumerate(zip(flat, shapes)):
        a = np.asarray(c, dtype=np.float64).reshape(-1)
        if a.size != int(np.prod(shp)):
            raise L.DimensionMismatch(f"{what} core {k}: {a.size} entries, expected {shp}")
        cores.append(a.reshape(shp, order="F"))
    return cores


def json_to_mps(x) -> TTvector:
    """``json_to_mps`` (``src/tt_tools.jl:765-774``): a TTvector from its JSON
    object (``N``, ``ttv_dims``, ``ttv_rks``, ``ttv_ot`` and ``ttv_vec``, each core
    a flat column-major list of shape ``(n_k, r_k, r_{k+1})``), as written by
    ``JSON3.write`` of a TTvector.  ``x`` is the parsed object or its text."""
    x = _json_obj(x)
    N = int(x["N"])
    dims = [int(v) for v in x["ttv_dims"]]
    rks = [int(v) for v in x["ttv_rks"]]
    ot = [int(v) for v in x["ttv_ot"]]
    if len(dims) != N or len(rks) != N + 1 or len(ot) != N or len(x["ttv_vec"]) != N:
        raise L.DimensionMismatch("inconsistent N, ttv_dims, ttv_rks, ttv_ot, ttv_vec")
    cores = _json_cores(x["ttv_vec"], [(dims[k], rks[k], rks[k + 1]) for k in range(N)], "ttv_vec")
    return TTvector(N, cores, dims, rks, ot)


def json_to_mpo(x) -> TToperator:
    """``json_to_mpo`` (``src/tt_tools.jl:776-785``): a TToperator from its JSON
    object (``tto_*`` fields, cores of shape ``(n_k, n_k, r_k, r_{k+1})``)."""
    x = _json_obj(x)
    N = int(x["N"])
    dims = [int(v) for v in x["tto_dims"]]
    rks = [int(v) for v in x["tto_rks"]]
    ot = [int(v) for v in x["tto_ot"]]
    if len(dims) != N or len(rks) != N + 1 or len(ot) != N or len(x["tto_vec"]) != N:
        raise L.DimensionMismatch("inconsistent N, tto_dims, tto_rks, tto_ot, tto_vec")
    cores = _json_cores(x["tto_vec"], [(dims[k], dims[k], rks[k], rks[k + 1]) for k in range(N)], "tto_vec")
    return TToperator(N, cores, dims, rks, ot)


def mps_to_json(x: TTvector) -> dict:
    """The JSON object ``JSON3.write`` makes of a TTvector (fields in struct
    order, cores flattened column-major): the inverse of ``json_to_mps``."""
    return {"N": x.N, "ttv_vec": [np.asarray(c).reshape(-1, order="F").tolist() for c in x.ttv_vec],
            "ttv_dims": list(x.ttv_dims), "ttv_rks": list(x.ttv_rks), "ttv_ot": list(x.ttv_ot)}


def mpo_to_json(A: TToperator) -> dict:
    """The JSON object of a TToperator: the inverse of ``json_to_mpo``."""
    return {"N": A.N, "tto_vec": [np.asarray(c).reshape(-1, order="F").tolist() for c in A.tto_vec],
            "tto_dims": list(A.tto_dims), "tto_rks": list(A.tto_rks), "tto_ot": list(A.tto_ot)}


def zeros_tt(dims, rks, ot=None) -> TTvector:
    d = len(dims)
    return TTvector(d, [np.zeros((dims[i], rks[i], rks[i + 1])) for i in range(d)], dims, rks, ot)


def rand_tt(dims, rks, seed: int) -> TTvector:
    """Seeded N(0,1) TT (drawn on the device with the library's Philox stream)."""
    dev = DeviceTT.randn(dims, rks, seed)
    out = dev.to_host()
    dev.free()
    return out


def tt_scale(a, x: TTvector) -> TTvector:
    """``*(a::Number, A::TTvector)`` (``src/tt_operations.jl:428-441``): scales the
    first core with ``ot == 0`` (host arithmetic on one core; no GPU work)."""
    if a == 0:
        return zeros_tt(x.ttv_dims, [1] * (x.N + 1))
    if 0 not in x.ttv_ot:
        raise ValueError("no core with ot == 0 to scale")
    i = x.ttv_ot.index(0)
    vec = list(x.ttv_vec)
    vec[i] = a * vec[i]
    return TTvector(x.N, vec, x.ttv_dims, list(x.ttv_rks), list(x.ttv_ot))


def _host_lincomb(coeffs, xs, plain_sum=False) -> TTvector:
    devs = [x._dev() for x in xs]
    out = lincomb_device(coeffs, devs)
    res = out.to_host()
    for d in devs + [out]:
        d.free()
    return res


def lincomb(A: Sequence[TTvector], x: Sequence[float]) -> TTvector:
    """``*(A::Vector{TTvector}, x::Vector)`` (``src/tt_operations.jl:252-258``)."""
    return _host_lincomb(list(x), list(A))


def is_leftorthogonal(x) -> bool:
    return all(o == 1 for o in x.ttv_ot[: x.N - 1])


def is_rightorthogonal(x) -> bool:
    return all(o == -1 for o in x.ttv_ot[1:])


def orthogonalize(x: TTvector, i: int = 1) -> TTvector:
    """``orthogonalize(x; i=1)`` (``src/tt_rounding.jl:65-97``)."""
    dev = x._dev()
    out = orthogonalize_batch([dev], i)[0]
    res = out.to_host()
    dev.free()
    out.free()
    return res


def norm(x: TTvector) -> float:
    """``norm(::TTvector)`` (``src/tt_rounding.jl:107-114``)."""
    dev = x._dev()
    v = dev.norm()
    dev.free()
    return v


def dot(A: TTvector, B: TTvector) -> float:
    """``dot(A, B)`` (``src/tt_operations.jl:302-321``)."""
    if A.ttv_dims != B.ttv_dims:
        raise L.DimensionMismatch("TT dimensions are not compatible")
    a, b = A._dev(), B._dev()
    v = float(dot_batch([a], [b])[0])
    a.free()
    b.free()
    return v


def dot_par(A: TTvector, B: TTvector) -> float:
    """``dot_par(A, B)`` (``src/tt_operations.jl:339-370``): the same inner
    product; the device evaluates every core's contraction in parallel already."""
    return dot(A, B)


def dot_operator(psi: TTvector, H: TToperator, phi: TTvector) -> float:
    """``dot_operator(psi, H, phi)`` = ``<psi|H|phi>`` without forming ``H phi``
    (``src/tt_operations.jl:385-421``)."""
    if psi.ttv_dims != phi.ttv_dims or tuple(H.tto_dims) != tuple(phi.ttv_dims):
        raise L.DimensionMismatch("TT dimensions are not compatible")
    a, b = psi._dev(), phi._dev()
    op = DeviceTTO(H, a.handle)
    v = float(dot_operator_batch([a], [op], [b])[0])
    for t in (a, b, op):
        t.free()
    return v


def partial_contraction(A: TTvector, B: TTvector, reverse: bool = True) -> List[np.ndarray]:
    """``partial_contraction(A, B; reverse)`` (``src/sketches.jl:1047-1070``):
    ``W[k]`` of shape ``(A.ttv_rks[k], B.ttv_rks[k])`` for ``k = 0..N``."""
    if A.ttv_dims != B.ttv_dims:
        raise L.DimensionMismatch("TT dimensions are not compatible")
    a, b = A._dev(), B._dev()
    h = a.handle
    Ws = [np.zeros((A.ttv_rks[k], B.ttv_rks[k]), order="F") for k in range(A.N + 1)]
    arr = (L.f64p * (A.N + 1))(*[L.ptr_f64(w) for w in Ws])
    h.check(h.lib.ttr_partial_contraction(h.h, a.id, b.id, 1 if reverse else 0, arr))
    a.free()
    b.free()
    return Ws


def tt_rounding(x, tol: float = 1e-12, rmax: int = 2 ** 14, direction: str = "left"):
    """``tt_rounding(x; tol, rmax, direction)`` (``src/tt_rounding.jl:146-154``);
    for a ``TToperator`` (``:275-277``) the operator is rounded as the TT of its
    ``(n^2, R, R')`` cores."""
    if isinstance(x, TToperator):
        v = TTvector(x.N, [np.reshape(c, (c.shape[0] * c.shape[1], c.shape[2], c.shape[3]), order="F") for c in x.tto_vec],
                     tuple(n * n for n in x.tto_dims), x.tto_rks, x.tto_ot)
        r = tt_rounding(v, tol=tol, rmax=rmax, direction=direction)
        return TToperator(r.N, [np.reshape(c, (n, n, c.shape[1], c.shape[2]), order="F") for c, n in zip(r.ttv_vec, x.tto_dims)],
                          x.tto_dims, r.ttv_rks, r.ttv_ot)
    dev = x._dev()
    out = round_svd_batch([dev], tol, rmax, direction)[0]
    res = out.to_host()
    dev.free()
    out.free()
    return res


def tt_rounding_(x: TTvector, tol: float = 1e-12, rmax: int = 2 ** 14, direction: str = "left") -> None:
    """``tt_rounding!`` (``src/tt_rounding.jl:158-167``): when ``x`` is already
    orthogonal in the sweep's direction its cores, ranks and orthogonality are
    replaced by the rounding; otherwise the reference rounds an orthogonalised copy
    and ``x`` keeps its values."""
    orth = is_leftorthogonal(x) if direction == "left" else is_rightorthogonal(x)
    r = tt_rounding(x, tol, rmax, direction)
    if orth:
        x.ttv_vec, x.ttv_rks, x.ttv_ot = r.ttv_vec, r.ttv_rks, r.ttv_ot


def default_rank_heuristic(y) -> List[int]:
    """``default_rank_heuristic`` (``src/tt_randtools.jl:31-50``)."""
    if isinstance(y, TTvector):
        rks, N = list(y.ttv_rks), y.N
    else:
        rks, N = list(y[0].ttv_rks), y[0].N
        for t in y[1:]:
            rks = [max(a, b) for a, b in zip(rks, t.ttv_rks)]
    ovs = int(math.ceil(0.5 * max(rks)))
    for k in range(1, N):
        rks[k] += ovs
    return rks


def _rks_arg(rks, N):
    if isinstance(rks, (int, np.integer)):
        r = [1] * (N + 1)
        for k in range(1, N):
            r[k] = int(rks)
        return r
    return [int(v) for v in rks]


def ttrand_rounding(*args, orthogonal: bool = True, seed: int = 1234, block_rks: Optional[int] = None,
                    sketch_prec: str = "f64", blocks=None, sketch_accum: str = "sketch") -> TTvector:
    """Randomized rounding (``src/tt_randtools.jl:112-361``).

    Forms: ``ttrand_rounding(y)``, ``ttrand_rounding(y, rks_or_rmax)``,
    ``ttrand_rounding(alpha, ys[, rks_or_rmax])``,
    ``ttrand_rounding(alpha, A, ys[, rks_or_rmax])`` (``alpha_1 A y_1 +
    sum_{j>1} alpha_j y_j``, the product never formed) and
    ``ttrand_rounding(A, y, b[, rks_or_rmax])`` (``A y - b``).  The default
    ranks follow ``default_rank_heuristic`` of ``y`` (or of ``ys``).  The sketch
    draws come from the library's counter-based Philox stream keyed by ``seed``;
    ``sketch_prec`` selects the precision of the sketch contractions (the QR and
    projections are FP64).  ``blocks`` (optional, per core) are sketch blocks as
    ``generate_sketch_blocks`` returns them, used instead of the library's draws."""
    op = None
    if isinstance(args[0], tuple):
        # NTuple form (src/tt_randtools.jl:364-466): the entrywise product of the
        # factors, never formed by the reference; target ranks are required
        # (the reference defines no default heuristic for a tuple)
        if len(args) < 2:
            raise TypeError("ttrand_rounding(y::Tuple, rks_or_rmax): ranks are required")
        if len(args[0]) == 1:
            return ttrand_rounding(args[0][0], args[1], orthogonal=orthogonal, seed=seed, block_rks=block_rks,
                                   sketch_prec=sketch_prec, blocks=blocks, sketch_accum=sketch_accum)
        ys = list(args[0])
        if any(not isinstance(t, TTvector) for t in ys):
            raise TypeError("a tuple argument must hold TTvectors")
        if any(t.ttv_dims != ys[0].ttv_dims for t in ys):
            raise L.DimensionMismatch("all TTs of the product must share dims")
        rks = _rks_arg(args[1], ys[0].N)
        devs = [t._dev() for t in ys]
        try:
            out = round_rand_tuple_batch([(devs, rks)], orthogonal=orthogonal, seed=seed, block_rks=block_rks,
                                         sketch_prec=sketch_prec, blocks=blocks, sketch_accum=sketch_accum)[0]
        finally:
            for d in devs:
                d.free()
        res = out.to_host()
        out.free()
        return res
    if isinstance(args[0], TTvector):
        y = args[0]
        rks = _rks_arg(args[1], y.N) if len(args) > 1 else default_rank_heuristic(y)
        alpha, ys = [1.0], [y]
    elif isinstance(args[0], TToperator):
        op, y, b = args[0], args[1], args[2]
        if y.ttv_dims != b.ttv_dims or tuple(op.tto_dims) != tuple(y.ttv_dims):
            raise L.DimensionMismatch("A, y and b must share dims")
        rks = _rks_arg(args[3], y.N) if len(args) > 3 else default_rank_heuristic(y)
        alpha, ys = [1.0, -1.0], [y, b]
    else:
        alpha = list(args[0])
        if isinstance(args[1], TToperator):
            op, ys, rest = args[1], list(args[2]), args[3:]
        else:
            ys, rest = list(args[1]), args[2:]
        if len(alpha) != len(ys) or any(t.ttv_dims != ys[0].ttv_dims for t in ys):
            raise L.DimensionMismatch("length(alpha) must equal length(y) and all dims must agree")
        if op is not None and tuple(op.tto_dims) != tuple(ys[0].ttv_dims):
            raise L.DimensionMismatch("operator and TT dims differ")
        rks = _rks_arg(rest[0], ys[0].N) if rest else default_rank_heuristic(ys)
    devs = [t._dev() for t in ys]
    dop = DeviceTTO(op) if op is not None else None
    item = (alpha, dop, devs, rks) if dop is not None else (alpha, devs, rks)
    out = round_rand_batch([item], orthogonal=orthogonal, seed=seed, block_rks=block_rks,
                           sketch_prec=sketch_prec, blocks=blocks, sketch_accum=sketch_accum)[0]
    res = out.to_host()
    for d in devs + [out] + ([dop] if dop is not None else []):
        d.free()
    return res


def tt_recursive_sketch(*args, orthogonal: bool = True, reverse: bool = True, seed: int = 1234,
                        block_rks: Optional[int] = None, p=None, oversampling: float = 1.0,
                        sketch_prec: str = "f64", blocks=None):
    """``tt_recursive_sketch([H,] A, rks; reverse=true, ...)`` (``src/sketches.jl:175-315``;
    operator form ``:317-472``, the product ``H A`` never formed; NTuple form
    ``:474-615``, the entrywise product of the tuple's TTs).

    Returns ``(W, sketch_rks)`` with ``W[k]`` of shape ``(r_k, sketch_rks[k])``,
    ``(r_k, R_k, sketch_rks[k])`` for the operator form, or
    ``(r_1k, ..., r_Mk, sketch_rks[k])`` for a tuple of M TTs."""
    if isinstance(args[0], tuple):
        if len(args[0]) > 1:  # factored: the product is never formed
            N = args[0][0].N
            rks = args[1]
            if isinstance(rks, (int, np.integer)):
                rks = [int(rks)] * (N + 1)
                rks[N if reverse else 0] = 1
            return _tuple_sketch(args[0], [int(v) for v in rks], orthogonal, seed, block_rks, p, oversampling,
                                 sketch_prec, blocks, forward=not reverse)
        P = _Product(args[0])
        W, srk = tt_recursive_sketch(P, *args[1:], orthogonal=orthogonal, reverse=reverse, seed=seed,
                                     block_rks=block_rks, p=p, oversampling=oversampling, sketch_prec=sketch_prec)
        return [P.split(w, k) for k, w in enumerate(W)], srk
    if isinstance(args[0], TToperator):
        H, A, rks = args
    else:
        H, (A, rks) = None, args
    if isinstance(rks, (int, np.integer)):
        r = [int(rks)] * (A.N + 1)
        r[A.N if reverse else 0] = 1
        rks = r
    if not reverse:
        return _forward_sketch(A, rks, orthogonal, seed, block_rks, p, oversampling, blocks, H)
    dev = A._dev()
    h = dev.handle
    dop = DeviceTTO(H, h) if H is not None else None
    o, keep = _opts(orthogonal, sketch_prec, block_rks, seed, p, oversampling, blocks=blocks)
    ra = L.as_i64(rks)
    bv, pp = block_structure(A.ttv_dims, rks, block_rks, p, oversampling)
    srk = [bv[k] * pp[k] for k in range(A.N + 1)]
    Hr = list(H.tto_rks) if H is not None else [1] * (A.N + 1)
    Ws = [np.zeros((A.ttv_rks[k] * Hr[k], srk[k]), order="F") for k in range(A.N + 1)]
    arr = (L.f64p * (A.N + 1))(*[L.ptr_f64(w) for w in Ws])
    out_srk = np.zeros(A.N + 1, np.int64)
    if dop is None:
        h.check(h.lib.ttr_sketch(h.h, dev.id, L.ptr_i64(ra), C.byref(o), arr, L.ptr_i64(out_srk)))
    else:
        h.check(h.lib.ttr_sketch_op(h.h, dop.id, dev.id, L.ptr_i64(ra), C.byref(o), arr, L.ptr_i64(out_srk)))
        dop.free()
    dev.free()
    assert [int(v) for v in out_srk] == srk
    if H is not None:
        Ws = [w.reshape((A.ttv_rks[k], Hr[k], srk[k]), order="F") for k, w in enumerate(Ws)]
    return Ws, srk


def _split_op(args):
    if isinstance(args[0], TToperator):
        return args[0], args[1:]
    return None, args


def tt_sketch(*args, reverse: bool = True, orthogonal: bool = True, seed: int = 1234,
              block_rks: Optional[int] = None, sketch_prec: str = "f64"):
    """``tt_sketch([H,] A, s; reverse=true, orthogonal, seed, block_rks)``
    (``src/sketches.jl:645-727``; operator form ``:729-817``, ``H A`` never formed).

    Returns ``(sketch, s_actual)``: the boundary sketch of length
    ``s_actual = bv_1 * cld(s, bv_1)``.  ``reverse=False`` (``:679-694,796-811,
    894-900``, FP64): the forward recursive sketch with ``cld(s, bv_{N+1})`` blocks at
    every core (the same draws), ``W_{N+1}`` returned."""
    H, (A, s) = _split_op(args)
    if not reverse:
        return _tt_sketch_forward(H, A, int(s), orthogonal, seed, block_rks)
    if isinstance(A, tuple) and len(A) > 1:
        # NTuple form (src/sketches.jl:819-909): the recursive sketch of the
        # product with p = cld(s, bv_1) blocks at every core, never formed
        if H is not None:
            raise TypeError("no operator form for a tuple of TTs")
        N = A[0].N
        bv, _ = block_structure(A[0].ttv_dims, [1] * (N + 1), block_rks)
        pc = -(-int(s) // bv[0])
        W, _ = _tuple_sketch(A, [1] * (N + 1), orthogonal, seed, block_rks, [pc] * N + [1], 1.0, sketch_prec)
        return np.reshape(W[0], -1, order="F").copy(), bv[0] * pc
    if isinstance(A, tuple):
        A = A[0]
    if H is not None and tuple(H.tto_dims) != tuple(A.ttv_dims):
        raise L.DimensionMismatch("operator and TT dims differ")
    dev = A._dev()
    dop = DeviceTTO(H, dev.handle) if H is not None else None
    try:
        out, sa = tt_sketch_batch([dev], s, [dop] if dop is not None else None, orthogonal, seed, block_rks,
                                  sketch_prec)
    finally:
        dev.free()
        if dop is not None:
            dop.free()
    return out[:, 0].copy(), sa


def _tt_sketch_forward(H, A, s, orthogonal, seed, block_rks):
    if isinstance(A, tuple) and H is not None:
        raise TypeError("no operator form for a tuple of TTs")
    if isinstance(A, tuple) and len(A) == 1:
        A = A[0]
    first = A[0] if isinstance(A, tuple) else A
    N = first.N
    bv, _ = block_structure_forward(first.ttv_dims, [1] * (N + 1), block_rks)
    pc = -(-int(s) // bv[N])
    pre = [H] if H is not None else []
    W, srk = tt_recursive_sketch(*pre, A, [1] * (N + 1), reverse=False, orthogonal=orthogonal, seed=seed,
                                 block_rks=block_rks, p=[1] + [pc] * N)
    return np.reshape(W[N], -1, order="F").copy(), bv[N] * pc


def tt_combined_sketch(*args, reverse: bool = True, orthogonal: bool = True, seed: int = 1234,
                       block_rks: Optional[int] = None, sketch_prec: str = "f64"):
    """``tt_combined_sketch([H,] A, rks_or_rmax, s; reverse=true, ...)``
    (``src/sketches.jl:930-1001``): the recursive sketch with the boundary rank
    raised to ``max(rks_1, min(s, rks_2))``, its boundary sketch extended to at
    least ``s`` entries by ``tt_sketch`` with ``seed + 1`` (each part weighted
    by ``sqrt(len/s)``).  Returns ``(sketch, W, sketch_rks)`` where ``W`` is the
    recursive sketch with ``W[1]`` replaced by ones and ``sketch_rks[1] = 1``.
    ``reverse=False`` (FP64): the same at the right boundary ``W[N+1]`` with the
    forward sketches."""
    H, (A, rks, s) = _split_op(args)
    M = len(A) if isinstance(A, tuple) else 0  # NTuple form (src/sketches.jl:1003-1040)
    if M and H is not None:
        raise TypeError("no operator form for a tuple of TTs")
    N = A[0].N if M else A.N
    b, inner = (0, 1) if reverse else (N, N - 1)
    if isinstance(rks, (int, np.integer)):
        rks = [int(rks)] * (N + 1)
        rks[N if reverse else 0] = 1
    rks = [int(v) for v in rks]
    rks[b] = max(rks[b], min(int(s), rks[inner]))
    pre = [H] if H is not None else []
    kw = dict(orthogonal=orthogonal, block_rks=block_rks, sketch_prec=sketch_prec)
    if not reverse:
        kw.pop("sketch_prec")
    W, srk = tt_recursive_sketch(*pre, A, rks, reverse=reverse, seed=seed, **kw)
    sketch = np.reshape(W[b], -1, order="F").copy()
    s_rec = srk[b]
    if s > s_rec:
        extra, _ = tt_sketch(*pre, A, int(s) - s_rec, seed=seed + 1, reverse=reverse, **kw)
        s = s_rec + len(extra)
        sketch = np.concatenate([math.sqrt(len(sketch) / s) * sketch, math.sqrt(len(extra) / s) * extra])
    W[b] = np.ones((1,) * (M + 1)) if M else (np.ones((1, 1)) if H is None else np.ones((1, 1, 1)))
    srk[b] = 1
    return sketch, W, srk


def _forward_sketch(A, rks, orthogonal, seed, block_rks, p, oversampling, blocks=None, H=None):
    """``tt_recursive_sketch([H,] A, rks; reverse=false)`` (``src/sketches.jl:241-303``;
    operator form ``:390-459``, which applies no oversampling), FP64.  The operator
    form returns ``W[k]`` of shape ``(r_k, R_k, s_k)``."""
    dev = A._dev()
    h = dev.handle
    if H is not None:
        oversampling = 1.0
    dop = DeviceTTO(H, h) if H is not None else None
    o, keep = _opts(orthogonal, "f64", block_rks, seed, p, oversampling, forward=True, blocks=blocks)
    ra = L.as_i64(rks)
    bv, pp = block_structure_forward(A.ttv_dims, rks, block_rks, p, oversampling)
    srk = [bv[k] * pp[k] for k in range(A.N + 1)]
    Hr = list(H.tto_rks) if H is not None else [1] * (A.N + 1)
    Ws = [np.zeros((A.ttv_rks[k] * Hr[k], srk[k]), order="F") for k in range(A.N + 1)]
    arr = (L.f64p * (A.N + 1))(*[L.ptr_f64(w) for w in Ws])
    out_srk = np.zeros(A.N + 1, np.int64)
    try:
        if dop is None:
            h.check(h.lib.ttr_sketch(h.h, dev.id, L.ptr_i64(ra), C.byref(o), arr, L.ptr_i64(out_srk)))
        else:
            h.check(h.lib.ttr_sketch_op(h.h, dop.id, dev.id, L.ptr_i64(ra), C.byref(o), arr, L.ptr_i64(out_srk)))
    finally:
        dev.free()
        if dop is not None:
            dop.free()
    assert [int(v) for v in out_srk] == srk
    if H is not None:
        Ws = [np.reshape(w, (A.ttv_rks[k], Hr[k], srk[k]), order="F") for k, w in enumerate(Ws)]
    return Ws, srk


def _rks_stta(rks, N):
    if isinstance(rks, (int, np.integer)):
        r = [1] * (N + 1)
        for k in range(1, N):
            r[k] = int(rks)
        return r
    return [int(v) for v in rks]


def stta_sketch(x: TTvector, rks, seed_left: int = 1234, seed_right: int = 5678, orthogonal: bool = True,
                block_rks: Optional[int] = None, blocks_left=None, blocks_right=None):
    """``stta_sketch(x, rks; seed_left, seed_right, orthogonal, block_rks)``
    (``src/sketches.jl:1126-1156``).  Returns ``(Omega, Psi)``: ``Omega[k]`` of shape
    ``(sl_{k+1}, sr_{k+1})`` and ``Psi[k]`` of shape ``(n_k, sl_k, sr_{k+1})``.
    ``stta_sketch(x, L, R)`` with explicit sketch TTs (``:1112-1124``) is the same
    pair for the partial contractions of (x, L) and (x, R)."""
    if isinstance(rks, TTvector):
        return _stta_sketch_tt(x, rks, seed_left)
    N = x.N
    rks = _rks_stta(rks, N)
    bvl, pl = block_structure_forward(x.ttv_dims, rks, block_rks, None, 1.5)
    bvr, pr = block_structure(x.ttv_dims, rks, block_rks)
    sl = [bvl[k] * pl[k] for k in range(N + 1)]
    sr = [bvr[k] * pr[k] for k in range(N + 1)]
    Om = [np.zeros((sl[k + 1], sr[k + 1]), order="F") for k in range(N - 1)]
    Ps = [np.zeros((x.ttv_dims[k], sl[k], sr[k + 1]), order="F") for k in range(N)]
    dev = x._dev()
    h = dev.handle
    o, keep = _opts(orthogonal, "f64", block_rks, seed_left, blocks=blocks_left, blocks_right=blocks_right)
    ra = L.as_i64(rks)
    om_arr = (L.f64p * max(N - 1, 1))(*[L.ptr_f64(w) for w in Om])
    ps_arr = (L.f64p * N)(*[L.ptr_f64(w) for w in Ps])
    osl, osr = np.zeros(N + 1, np.int64), np.zeros(N + 1, np.int64)
    try:
        h.check(h.lib.ttr_stta_sketch(h.h, dev.id, L.ptr_i64(ra), C.byref(o), int(seed_right) & 0xFFFFFFFFFFFFFFFF,
                                      om_arr, ps_arr, L.ptr_i64(osl), L.ptr_i64(osr)))
    finally:
        dev.free()
    assert [int(v) for v in osl] == sl and [int(v) for v in osr] == sr
    return Om, Ps


def _stta_sketch_tt(x: TTvector, Lt: TTvector, Rt: TTvector):
    if x.ttv_dims != Lt.ttv_dims or x.ttv_dims != Rt.ttv_dims:
        raise L.DimensionMismatch("TT dimensions are not compatible")
    N = x.N
    Om = [np.zeros((Lt.ttv_rks[k + 1], Rt.ttv_rks[k + 1]), order="F") for k in range(N - 1)]
    Ps = [np.zeros((x.ttv_dims[k], Lt.ttv_rks[k], Rt.ttv_rks[k + 1]), order="F") for k in range(N)]
    devs = [t._dev() for t in (x, Lt, Rt)]
    h = devs[0].handle
    om_arr = (L.f64p * max(N - 1, 1))(*[L.ptr_f64(w) for w in Om])
    ps_arr = (L.f64p * N)(*[L.ptr_f64(w) for w in Ps])
    try:
        h.check(h.lib.ttr_stta_sketch_tt(h.h, devs[0].id, devs[1].id, devs[2].id, om_arr, ps_arr))
    finally:
        for d in devs:
            d.free()
    return Om, Ps


def stta_batch(items, seed_left: int = 1234, seed_right: int = 5678, orthogonal: bool = True,
               block_rks: Optional[int] = None, blocks_left=None, blocks_right=None) -> List[DeviceTT]:
    """Batched ``stta``: ``items`` is a list of ``(alphas, [DeviceTT...], rks)``."""
    h = items[0][1][0].handle
    keep = []
    arr = (L.RandItem * len(items))()
    for i, (alpha, terms, rks) in enumerate(items):
        a = np.ascontiguousarray(np.asarray(alpha, dtype=np.float64))
        ids = L.as_i64([t.id for t in terms])
        r = L.as_i64(_rks_stta(rks, terms[0].d))
        keep += [a, ids, r]
        arr[i].nterms = len(terms)
        arr[i].alpha = L.ptr_f64(a)
        arr[i].term_ids = L.ptr_i64(ids)
        arr[i].op_id = -1
        arr[i].rks = L.ptr_i64(r)
    o, k2 = _opts(orthogonal, "f64", block_rks, seed_left, blocks=blocks_left, blocks_right=blocks_right)
    out = np.zeros(len(items), np.int64)
    h.check(h.lib.ttr_stta(h.h, len(items), arr, C.byref(o), int(seed_right) & 0xFFFFFFFFFFFFFFFF, L.ptr_i64(out)))
    return [DeviceTT(h, t) for t in out]


def stta(*args, seed_left: int = 1234, seed_right: int = 5678, orthogonal: bool = True,
         block_rks: Optional[int] = None, blocks_left=None, blocks_right=None) -> TTvector:
    """Streaming TT approximation (``src/tt_randtools.jl:552-638``).

    Forms: ``stta(y[, rks_or_rmax])`` and ``stta(alpha, ys[, rks_or_rmax])``; the
    default ranks follow ``default_rank_heuristic`` of ``y`` (or ``ys``).  The
    left sketch draws from ``seed_left``, the right one from ``seed_right``
    (the library's counter-based streams)."""
    if isinstance(args[0], TTvector):
        y = args[0]
        alpha, ys = [1.0], [y]
        rks = _rks_stta(args[1], y.N) if len(args) > 1 else default_rank_heuristic(y)
    else:
        alpha, ys = list(args[0]), list(args[1])
        if len(alpha) != len(ys) or any(t.ttv_dims != ys[0].ttv_dims for t in ys):
            raise L.DimensionMismatch("length(alpha) must equal length(y) and all dims must agree")
        rks = _rks_stta(args[2], ys[0].N) if len(args) > 2 else default_rank_heuristic(ys)
    devs = [t._dev() for t in ys]
    try:
        out = stta_batch([(alpha, devs, rks)], seed_left, seed_right, orthogonal, block_rks, blocks_left,
                         blocks_right)[0]
        res = out.to_host()
        out.free()
    finally:
        for d in devs:
            d.free()
    return res


# ------------------------------------------- sketched Krylov basis and Rayleigh-Ritz

def _first_root(x: TTvector) -> int:
    return list(x.ttv_ot).index(0)


def expand_basis_(H: TToperator, B_window: List[TTvector], W_B_window: list, B_sketch_window: List[np.ndarray],
                  rmax: int, s: int, orthogonal: bool = True, block_rks: Optional[int] = 8, seed: int = 1234,
                  k_trunc: Optional[int] = None):
    """``expand_basis!`` (``src/tt_solvers.jl:354-455``): one sketch-orthogonal
    Arnoldi step over the sliding window, which it mutates.

    The boundary sketch of ``H b_prev`` (``tt_combined_sketch``, ranks ``2 rmax``,
    size ``s``) gives ``h``, the least-squares coefficients against the window's
    boundary sketches.  The reference's left-to-right sweep over the terms
    ``H b_prev, -h_i b_i`` with the sketches of the same seed and ranks is the
    operator form of the randomized rounding, so it runs as one
    ``ttrand_rounding([1, -h], H, [b_prev, window...], [2 rmax, ..., 1])`` on the
    device (which sketches the terms itself: the same draws, so the same sketches
    as the window's stored ones up to rounding); then ``tt_rounding(rmax)``, the
    combined sketch of the result and the normalisation by its norm.  Returns
    ``(b_new, sketch_b_new, sketch_Hb_prev)``."""
    if k_trunc is None:
        k_trunc = len(B_window)
    b_prev = B_window[-1]
    N = b_prev.N
    kw = dict(orthogonal=orthogonal, seed=seed, block_rks=block_rks)
    sketch_Hb_prev, _, _ = tt_combined_sketch(H, b_prev, 2 * rmax, s, **kw)
    h = np.linalg.lstsq(np.column_stack(B_sketch_window), sketch_Hb_prev, rcond=None)[0]
    rks = [2 * rmax] * (N + 1)
    rks[N] = 1
    y = ttrand_rounding([1.0] + [-float(v) for v in h], H, [b_prev] + list(B_window), rks, **kw)
    b_new = tt_rounding(y, rmax=rmax)
    sketch_b_new, W_B_new, _ = tt_combined_sketch(b_new, 2 * rmax, s, **kw)
    beta = float(np.linalg.norm(sketch_b_new))
    sketch_b_new = sketch_b_new / beta
    for i in range(1, _first_root(b_new) + 1):  # W_B_new[i] ./= beta, i = 2:findfirst(ot .== 0)
        W_B_new[i] = W_B_new[i] / beta
    b_new = b_new / beta
    for win, item in ((B_window, b_new), (B_sketch_window, sketch_b_new), (W_B_window, W_B_new)):
        win.append(item)
        if len(win) > k_trunc:
            win.pop(0)
    return b_new, sketch_b_new, sketch_Hb_prev


def _eig_sorted(M: np.ndarray, B: Optional[np.ndarray] = None):
    # Julia's eigen: eigenvalues sorted by (real, imag), vectors in the same order
    if B is None:
        lam, V = np.linalg.eig(M)
    else:
        import scipy.linalg
        lam, V = scipy.linalg.eig(M, B)
    order = sorted(range(len(lam)), key=lambda i: (lam[i].real, lam[i].imag))
    return lam[order], V[:, order]


def sketched_rr(H: TToperator, B: Sequence[TTvector], C: np.ndarray, D: np.ndarray, rmax: int,
                orthogonal: bool = True, block_rks: Optional[int] = 8, seed: int = 1234,
                kappa_max: Optional[float] = None, stable: bool = False, seed_samp: Optional[int] = None,
                verbose: bool = True):
    """Sketched Rayleigh-Ritz (``src/tt_solvers.jl:466-514``; Tropp & Nakatsukasa 2022).

    ``C = Omega B`` and ``D = Omega H B`` (``s x d``).  ``stable=False`` solves
    ``C M = D`` in the least-squares sense and takes ``eigen(M)``;
    ``stable=True`` whitens with the thin SVD of ``C`` (directions with
    ``sigma/sigma_1 <= 1/kappa_max`` dropped) and solves the generalized problem.
    The Ritz vector ``sum_j y_j B_j`` is randomized-rounded to ranks ``4 rmax``
    and SVD-rounded at ``tol = res_sketch^2/2`` on the device, then normalised
    (NaN when that rounding returned zeros, as the reference's ``1/0``); the
    sampled residual compares boundary sketches of ``H X`` and ``X`` of size ``2N``
    drawn with ``seed_samp`` (random when None, the reference's ``rand(Int)``).
    Returns ``(lambdas, X, res_sketch, res_sample)``."""
    N = B[0].N
    if kappa_max is None:
        kappa_max = 1.0 / math.sqrt(np.finfo(np.float64).eps)
    if stable:
        U, sig, Vt = np.linalg.svd(C, full_matrices=False)
        if verbose:
            print(f"Sketch condition number κ(C) ≈ {sig[0] / sig[-1]:.4g}")
        keep = [i for i in range(len(sig)) if sig[i] / sig[0] > 1.0 / kappa_max]
        k = keep[-1] + 1 if keep else 1
        if k < len(B) and verbose:
            print(f"Warning: basis truncated from d={len(B)} to k={k} (ill-conditioned directions removed)")
        Vk, sk = Vt[:k].T, sig[:k]
        lam, Z = _eig_sorted(U[:, :k].T @ D @ Vk, np.diag(sk))
        Y = np.real(Vk @ Z)
        for j in range(k):
            Y[:, j] = Y[:, j] / np.linalg.norm(C @ Y[:, j])
    else:
        lam, Y = _eig_sorted(np.linalg.lstsq(C, D, rcond=None)[0])
        Y = np.real(Y)
    lam = np.real(lam)
    res_sketch = float(np.linalg.norm(D @ Y[:, 0] - lam[0] * (C @ Y[:, 0])))
    y = Y[:, 0] / np.linalg.norm(Y[:, 0])
    X = ttrand_rounding([float(v) for v in y], list(B), 4 * rmax, orthogonal=orthogonal, seed=seed,
                        block_rks=block_rks)
    X = tt_rounding(X, tol=res_sketch ** 2 / 2)
    with np.errstate(divide="ignore", invalid="ignore"):
        X = tt_scale(np.float64(1.0) / np.float64(norm(X)), X)
    if seed_samp is None:
        seed_samp = int(np.random.default_rng().integers(-(1 << 63), (1 << 63) - 1))
    kw = dict(orthogonal=orthogonal, seed=seed_samp, block_rks=block_rks)
    S1, _ = tt_sketch(H, X, 2 * N, **kw)
    S2, _ = tt_sketch(X, 2 * N, **kw)
    with np.errstate(invalid="ignore"):
        res_sample = float(np.linalg.norm(S1 - lam[0] * S2) / np.linalg.norm(S2))
    return lam, X, res_sketch, res_sample


def sketched_rayleigh_ritz(H: TToperator, b0: TTvector, d, rmax: Optional[int] = None, orthogonal: bool = True,
                           block_rks: Optional[int] = 8, seed: int = 1234, sketch_size: Optional[int] = None,
                           k_trunc: Optional[int] = None, e_nuc: float = 0.0, validate_seed: bool = False,
                           verbose: bool = True, sample_seeds=None):
    """``sketched_rayleigh_ritz`` (``src/tt_solvers.jl:548-689``).

    ``sketched_rayleigh_ritz(H, b0, d, rmax)``: a ``d``-dimensional
    sketch-orthogonal Krylov basis from ``b0`` (normalised by its boundary sketch),
    a sketched Rayleigh-Ritz after every expansion and a final one over all ``d``
    vectors; returns ``(e_sketch, psi, history)`` with history entries
    ``dict(iter, e_sketch, e_true, res_sketch, res_sample)``.

    ``sketched_rayleigh_ritz(H, b0, schedule)`` with ``schedule`` a list of dicts
    (``d``, ``rmax`` and optionally ``seed``, ``sketch_size``, ``k_trunc``,
    ``label``): one stage per entry, each restarted from the previous stage's Ritz
    vector rounded to the stage's ``rmax``; returns ``(e, psi, stages,
    seeds_used)``.  ``sample_seeds`` (optional) fixes the seeds of the sampled
    residuals, which the reference draws with ``rand(Int)``."""
    if not isinstance(d, (int, np.integer)):
        return _srr_schedule(H, b0, list(d), orthogonal, block_rks, seed, e_nuc, validate_seed, verbose)
    d, rmax = int(d), int(rmax)
    N = b0.N
    if sketch_size is None:
        sketch_size = 4 * d
    if k_trunc is None:
        k_trunc = d - 1
    samp = list(sample_seeds) if sample_seeds is not None else [None] * d
    kw = dict(orthogonal=orthogonal, seed=seed, block_rks=block_rks)
    sk1, W_B_1, _ = tt_combined_sketch(b0, 2 * rmax, sketch_size, **kw)
    s = len(sk1)
    beta = float(np.linalg.norm(sk1))
    sk1 = sk1 / beta
    for i in range(1, _first_root(b0) + 1):
        W_B_1[i] = W_B_1[i] / beta
    B = [b0 / beta]
    sketch_b, sketch_Hb = [sk1], []
    B_window, W_B_window, B_sketch_window = [B[0]], [W_B_1], [sk1]
    history = []

    def srr(Bs, ss):
        C = np.column_stack(sketch_b[:len(Bs)])
        Dm = np.column_stack(sketch_Hb[:len(Bs)])
        lam, psi, r1, r2 = sketched_rr(H, Bs, C, Dm, rmax, seed_samp=ss, verbose=verbose, **kw)
        with np.errstate(invalid="ignore"):
            e_true = dot_operator(psi, H, psi) + e_nuc
        return lam[0] + e_nuc, psi, dict(iter=len(Bs), e_sketch=lam[0] + e_nuc, e_true=e_true,
                                         res_sketch=r1, res_sample=r2)

    for j in range(2, d + 1):
        if verbose:
            print(f"─── k = {j - 1} " + "─" * 44)
        b_new, sk_new, sk_Hb = expand_basis_(H, B_window, W_B_window, B_sketch_window, rmax, s, k_trunc=k_trunc,
                                             **kw)
        B.append(b_new)
        sketch_b.append(sk_new)
        sketch_Hb.append(sk_Hb)
        _, psi, rec = srr(B[:j - 1], samp[j - 2])
        history.append(rec)
        if verbose:
            print(f"  Sketched Ritz value:    {rec['e_sketch']}\n  True Ritz value:        {rec['e_true']}\n"
                  f"  Sketch residual:  {rec['res_sketch']}\n  Sampled residual: {rec['res_sample']}")
    if verbose:
        print(f"─── k = {d} (final) " + "─" * 45)
    skH, _, _ = tt_combined_sketch(H, B[-1], 2 * rmax, s, **kw)
    sketch_Hb.append(skH)
    e, psi, rec = srr(B, samp[d - 1])
    history.append(rec)
    if verbose:
        print(f"  Sketched Ritz:    {rec['e_sketch']}\n  True Ritz:        {rec['e_true']}\n"
              f"  Sketch residual:  {rec['res_sketch']}\n  Sampled residual: {rec['res_sample']}")
    return e, psi, history


def _srr_schedule(H, b0, schedule, orthogonal, block_rks, seed, e_nuc, validate_seed, verbose):
    # multi-stage form (src/tt_solvers.jl:622-689)
    stages, seeds_used = [], []
    psi, lam = b0, 0.0
    rng = np.random.default_rng()
    for i, stage in enumerate(schedule, start=1):
        d, rmax = int(stage["d"]), int(stage["rmax"])
        s = int(stage.get("seed", seed + i - 1))
        sk = int(stage.get("sketch_size", 4 * d))
        kt = int(stage.get("k_trunc", d - 1))
        lbl = stage.get("label", f"d={d}, rmax={rmax}")
        b_init = psi if i == 1 else tt_rounding(psi, rmax=rmax)
        if validate_seed:
            while abs(np.linalg.norm(tt_combined_sketch(b_init, 2 * rmax, sk, orthogonal=orthogonal, seed=s,
                                                         block_rks=block_rks)[0]) ** 2 - 1) > 0.1:
                s = int(rng.integers(-(1 << 63), (1 << 63) - 1))
            if verbose:
                print(f"Stage {i}: using seed {s}")
        seeds_used.append(s)
        lam, psi, hist = sketched_rayleigh_ritz(H, b_init, d, rmax, orthogonal=orthogonal, block_rks=block_rks,
                                                seed=s, k_trunc=kt, sketch_size=sk, e_nuc=e_nuc, verbose=verbose)
        stages.append((lbl, hist))
    return lam, psi, stages, seeds_used
