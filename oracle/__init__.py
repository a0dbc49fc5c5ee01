"""FP64 CPU oracle for the TT-rounding hot path.

TEST INFRASTRUCTURE ONLY.  This package is the checker the parity tests, the
smoke check and the CPU-baseline leg of the benchmark compare the HIP library
against.  It is never imported by the product package (``ttround``) and is not a
fallback for it.

It restates, in numpy on the host, the algorithms of the Julia package
``TensorTrains`` (reference files ``src/tt_rounding.jl``, ``src/tt_randtools.jl``,
``src/sketches.jl``, ``src/tt_tools.jl``, ``src/tt_operations.jl``); each function
cites the reference lines it follows.  Cores use the reference layout
``(n_k, r_{k-1}, r_k)`` and Julia's column-major reshapes (``order='F'``).

What pins each function (besides the reference's own tolerance tests, which the
test-suite restates) is listed in the docstring of ``tests/test_oracle.py``.
Parity unpinned: none of the functions here; the Julia RNG stream itself is
not restated (the sketch draws are always explicit inputs).
"""
from .tt_oracle import *  # noqa: F401,F403
from . import tt_oracle, philox  # noqa: F401
from .tt_oracle import _sketch_for  # noqa: F401
