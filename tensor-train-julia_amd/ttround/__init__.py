"""ttround: MI355X (gfx950) TT-rounding library, Python host mirror.

The numerical work runs in ``libttround.so`` (hand-written HIP kernels behind
the C-ABI of ``include/ttround.h``).  This package mirrors the TensorTrains
Julia API for the rounding path; see ``ttround.api``.
"""
from ._lib import (DimensionMismatch, NumericalFailure, DeviceError, Handle, default_handle, load,  # noqa: F401
                   LIB_PATH, SIGNATURES)
from .api import *  # noqa: F401,F403
from .api import (DeviceTT, DeviceTTO, TTvector, TToperator, default_rank_heuristic, lincomb, lincomb_device,  # noqa: F401
                  mpo_apply_device, norm, norm_batch, orthogonalize, orthogonalize_batch, r_and_d_to_rks, rand_tt,
                  round_rand_batch, round_svd_batch, sketch_blocks, sketch_draws, tt_recursive_sketch, tt_rounding,
                  tt_sketch, tt_sketch_batch, tt_combined_sketch, stta, stta_batch, stta_sketch,
                  block_structure_forward,
                  tt_rounding_, tt_scale, ttrand_rounding, zeros_tt, is_leftorthogonal, is_rightorthogonal,
                  dot, dot_par, dot_operator, dot_batch, dot_operator_batch, partial_contraction,
                  hadamard_device, round_rand_tuple_batch, json_to_mps, json_to_mpo, mps_to_json, mpo_to_json,
                  expand_basis_, sketched_rr, sketched_rayleigh_ritz)

__version__ = "0.1.0"
from .models import hubbard_1d_mpo  # noqa: F401,E402
