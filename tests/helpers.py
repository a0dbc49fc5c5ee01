"""Conversions between the oracle's TT objects and the host mirror's."""
import numpy as np

import oracle as O


def to_api(x):
    import ttround as T
    return T.TTvector(x.N, [c.copy() for c in x.cores], x.dims, list(x.rks), list(x.ot))


def to_oracle(x):
    return O.TT([np.asarray(c) for c in x.ttv_vec], x.ttv_dims, list(x.ttv_rks), list(x.ttv_ot))


def dev_to_oracle(dx):
    return O.TT([dx.core(k) for k in range(dx.d)], dx.dims, list(dx.rks), list(dx.ot))


def rel_err_dense(x, y):
    X, Y = O.ttv_to_tensor(x), O.ttv_to_tensor(y)
    return float(np.linalg.norm(X - Y) / np.linalg.norm(X))


def device_blocks_provider(dims, rks, **kw):
    """Callable giving the oracle the exact sketch blocks the device uses."""
    import ttround as T
    cache = {}

    def get(k):
        if k not in cache:
            cache[k] = T.sketch_blocks(dims, rks, k, **kw)
        return cache[k]
    return get
