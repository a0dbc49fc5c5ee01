"""C-ABI checks that need no GPU: the library loads and exports every symbol
declared in include/ttround.h; host-side integer logic equals the oracle's."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ttround.h")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ttr_[a-z_0-9]+)\s*\(", txt)))


def test_header_symbols_are_bound():
    import ttround
    assert set(declared_symbols()) == set(ttround.SIGNATURES)


def test_library_exports_every_symbol():
    import ttround
    lib = ttround.load()
    for name in declared_symbols():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", ttround.LIB_PATH], capture_output=True, text=True).stdout
    for name in declared_symbols():
        assert re.search(r"\bT " + name + r"\b", out), name
    assert lib.ttr_version().decode().startswith("libttround")


def test_host_integer_logic_matches_oracle():
    import ttround as T
    for dims, rks in [((4,) * 40, [256] * 41), ((2, 3, 4), [1, 5, 5, 1]), ((3,) * 50, [7] * 51)]:
        assert T.r_and_d_to_rks(rks, dims, rmax=max(rks)) == O.r_and_d_to_rks(rks, dims, rmax=max(rks))
    for dims, rks, br in [((4,) * 32, [1] + [144] * 31 + [1], 144), ((2,) * 6, [1, 9, 9, 9, 9, 9, 1], 2),
                          ((4,) * 40, [1] + [300] * 39 + [1], None)]:
        assert T.api.block_structure(dims, rks, br) == tuple(O.sketch_structure(dims, rks, br))


def test_create_without_gpu_reports_error():
    import ttround as T
    try:
        import torch
        if torch.cuda.device_count() > 0:
            pytest.skip("a GPU is visible")
    except ImportError:
        pass
    with pytest.raises(Exception):
        T.Handle(0)


def test_hubbard_builder_matches_oracle():
    """The product's Hubbard MPO (bench input) equals the oracle's, which is
    pinned against the dense fermionic Hamiltonian."""
    import ttround as T
    for L in (2, 3, 6):
        A = T.hubbard_1d_mpo(L, 1.0, 4.0)
        H = O.hubbard_1d_mpo(L, 1.0, 4.0)
        assert A.tto_rks == H.rks and A.tto_dims == H.dims
        for a, b in zip(A.tto_vec, H.cores):
            np.testing.assert_array_equal(a, b)
    # a single site is the on-site term alone
    A1 = T.hubbard_1d_mpo(1, 1.0, 4.0)
    np.testing.assert_array_equal(A1.tto_vec[0][:, :, 0, 0], np.diag([0.0, 0.0, 0.0, 4.0]))
