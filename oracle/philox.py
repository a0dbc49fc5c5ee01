"""Philox4x32-10 counter-based generator and the library's normal draws, in numpy.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Restates the published
algorithm (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as easy as
1, 2, 3", SC'11) and the draw mapping documented in the library: draw ``e`` of
core ``k`` is counter ``(e_lo, e_hi, k, tag)`` under key ``(seed_lo, seed_hi)``;
two 53-bit uniforms ``u1 = (a + 1/2) 2^-53``, ``u2 = b 2^-53``; normal
``sqrt(-2 log u1) cos(2 pi u2)``.  Pinned by the Random123 known-answer vectors
in tests/test_oracle.py.
"""
import numpy as np

M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
W0, W1 = np.uint64(0x9E3779B9), np.uint64(0xBB67AE85)
MASK = np.uint64(0xFFFFFFFF)
SKETCH_TAG = 0x5EED
INPUT_TAG = 0x7A11


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    c = [np.asarray(v, dtype=np.uint64) & MASK for v in (c0, c1, c2, c3)]
    k0 = np.asarray(k0, dtype=np.uint64) & MASK
    k1 = np.asarray(k1, dtype=np.uint64) & MASK
    for _ in range(10):
        p0 = M0 * c[0]
        p1 = M1 * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & MASK
        c = [hi1 ^ c[1] ^ k0, lo1, hi0 ^ c[3] ^ k1, lo0]
        k0 = (k0 + W0) & MASK
        k1 = (k1 + W1) & MASK
    return c


def normal_draws(seed: int, core: int, count: int, tag: int = SKETCH_TAG, start: int = 0) -> np.ndarray:
    e = np.arange(start, start + count, dtype=np.uint64)
    seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    x = philox4x32_10(e & MASK, e >> np.uint64(32), np.full_like(e, core), np.full_like(e, tag),
                      np.uint64(seed & 0xFFFFFFFF), np.uint64(seed >> 32))
    a = ((x[0] << np.uint64(32)) | x[1]) >> np.uint64(11)
    b = ((x[2] << np.uint64(32)) | x[3]) >> np.uint64(11)
    u1 = (a.astype(np.float64) + 0.5) * 2.0 ** -53
    u2 = b.astype(np.float64) * 2.0 ** -53
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2)


def rand_tt_cores(dims, rks, seed: int):
    """N(0,1) cores with the library's input stream: draw e (reference order
    i + n*(a + r*b)) of core k."""
    cores = []
    for k, n in enumerate(dims):
        cnt = n * rks[k] * rks[k + 1]
        g = normal_draws(seed, k, cnt, tag=INPUT_TAG)
        cores.append(np.reshape(g, (n, rks[k], rks[k + 1]), order="F"))
    return cores
