"""Input operators for the benchmarks and examples (host side, numpy).

``hubbard_1d_mpo`` builds the closed-form matrix product operator of the open
1D Hubbard chain, one site per core with local dimension 4 (basis index
n_up + 2 n_dn), the Jordan-Wigner strings folded into the hopping terms.  The
reference assembles the same Hamiltonian from ``hubbard_1D`` + ``hV_to_mpo``
(``src/models.jl:403-419,180-231``) on 2L spin-orbital sites; its interaction
enters as ``V[i]*(term + h.c.)``, i.e. ``2U n_up n_dn``.  This builder uses the
textbook ``U n_up n_dn`` (pass ``U=2U`` for the reference's Hamiltonian).
"""
import numpy as np

from .api import TToperator


def hubbard_1d_mpo(L: int, t: float = 1.0, U: float = 4.0) -> TToperator:
    """``H = -t sum_{i,s} (c+_{i,s} c_{i+1,s} + h.c.) + U sum_i n_{i,up} n_{i,dn}``
    as a TToperator of bond dimension 6 (cores ``[i_out, j_in, a, b]``)."""
    eye = np.eye(4)
    c_up = np.zeros((4, 4))
    c_up[0, 1] = c_up[2, 3] = 1.0          # removes the up electron (up mode first)
    c_dn = np.zeros((4, 4))
    c_dn[0, 2], c_dn[1, 3] = 1.0, -1.0     # passes the up mode: sign (-1)^{n_up}
    parity = np.diag([1.0, -1.0, -1.0, 1.0])
    n_updn = np.diag([0.0, 0.0, 0.0, 1.0])
    # operator-valued W[a, b]; H = W_1 ... W_L, first site row 5, last site column 0
    W = np.zeros((6, 6, 4, 4))
    W[0, 0] = eye
    W[1, 0], W[2, 0] = c_up, c_up.T
    W[3, 0], W[4, 0] = c_dn, c_dn.T
    W[5, 0] = U * n_updn
    W[5, 1] = -t * (c_up.T @ parity)
    W[5, 2] = -t * (parity @ c_up)
    W[5, 3] = -t * (c_dn.T @ parity)
    W[5, 4] = -t * (parity @ c_dn)
    W[5, 5] = eye
    cores = []
    for i in range(L):
        Wi = W[5:6] if i == 0 else (W[:, 0:1] if i == L - 1 else W)
        if L == 1:
            Wi = W[5:6, 0:1]
        cores.append(np.ascontiguousarray(np.transpose(Wi, (2, 3, 0, 1))))
    rks = [1] + [6] * (L - 1) + [1]
    return TToperator(L, cores, (4,) * L, rks)
