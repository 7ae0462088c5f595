"""Particle-removal spectral weights: CPU restatement (test infrastructure only).

The reference plans a spectral function but ships none (README.md:14-15 "coming soon",
modules/green_function.h:18,90 commented-out spectral_function_sampler), so there is no
reference algorithm to follow.  The product computes it as follows, and this module restates
it for the tests:

    Psi_L^{N+1}(z_1..z_N, z_0) = Psi_L^N(z_1..z_N) prod_j (z_j - z_0)^m.

The z_0^l coefficient of prod_j (z_j - z_0)^m is (-1)^l e_k(M), M the multiset {z_j, each m
times} and k = m N - l.  So c_l Psi_L^{N+1} is proportional to e_k(M) Psi_L^N, a state of the
N-particle sector dL = k.  Newton's identities give

    e_k(M) = sum_{lambda |- k} (-1)^(k - len(lambda)) / z_lambda  prod_i P_{lambda_i}(M),

with P_i(M) = m p_i and z_lambda = prod_i i^(m_i) m_i!.  With S-orthonormal generalized
eigenvectors v_j of (H, S) in that sector, the weight of eigenstate j is |v_j^H S a|^2 / (a^T S a).

Pinned by exact_lll.exact_removal (coefficient extraction from the expanded Laughlin
polynomial, tiny N) in tests/test_oracle.py.
"""
from math import factorial

import numpy as np


def newton_coefficients(dl, m, states):
    """a_lambda of e_dl(M) over power-sum monomials; states = list of (dL, pairs)."""
    out = []
    for dls, pairs in states:
        assert dls == dl
        pairs = [(k, r) for k, r in pairs if k > 0]     # (0, 1) is the dL = 0 state, 1
        length = sum(r for _, r in pairs)
        z = 1
        for k, r in pairs:
            z *= k ** r * factorial(r)
        out.append((-1) ** (dl - length) * m ** length / z)
    return np.array(out, dtype=np.float64)


def removal_weights(S, H, a):
    """(energies, weights) of the state a in the basis with metric S and Hamiltonian H."""
    from scipy.linalg import eigh
    S = (S + S.conj().T) / 2
    H = (H + H.conj().T) / 2
    ev, V = eigh(H, S)
    sa = S @ a
    w = np.abs(V.conj().T @ sa) ** 2 / float(np.real(a @ sa))
    return ev, w
