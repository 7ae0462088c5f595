"""Writes the small fixtures under tests/golden/ (run from the repository root):

  geneig_scipy.npz   three Hermitian pairs (S positive definite) with scipy.linalg.eigh
                     eigenvalues, the LAPACK anchor of the generalized eigensolver
  exact_n4_m3.npz    exact S and H of every sector dL = 0..4 for N = 4, m = 3,
                     V = 0.025 r^2 + 1e-4 r^4 (power sums of z / R), from the monomial
                     expansion in oracle/exact_lll.py, with the basis listing

The reference ships no test vectors; these are computed in this container."""
import os
import sys

import numpy as np
import scipy.linalg

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle.exact_lll import exact_blocks          # noqa: E402
from oracle.partitions import Basis                # noqa: E402


def geneig_fixture():
    rng = np.random.default_rng(20240611)
    out = {}
    for k, d in enumerate((3, 8, 20)):
        A = rng.standard_normal((d, 2 * d)) + 1j * rng.standard_normal((d, 2 * d))
        S = A @ A.conj().T / (2 * d) + 0.05 * np.eye(d)
        B = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
        H = (B + B.conj().T) / 2
        out[f"S{k}"] = S
        out[f"H{k}"] = H
        out[f"E{k}"] = scipy.linalg.eigh(H, S, eigvals_only=True)
    np.savez(os.path.join(HERE, "geneig_scipy.npz"), **out)


def exact_fixture():
    N, m, w, eps = 4, 3, 0.05, 1e-4
    out = {"N": N, "m": m, "w": w, "eps": eps}
    for dl in range(5):
        b = Basis(N, [dl])
        S, H = exact_blocks(N, m, b.states, w, eps)
        out[f"S{dl}"] = np.array(S)
        out[f"H{dl}"] = np.array(H)
        pairs = np.zeros((len(b.states), 8), dtype=np.int32)
        for i, (_, p) in enumerate(b.states):
            for q, (k, r) in enumerate(p):
                pairs[i, 2 * q], pairs[i, 2 * q + 1] = k, r
        out[f"pairs{dl}"] = pairs
    np.savez(os.path.join(HERE, "exact_n4_m3.npz"), **out)


if __name__ == "__main__":
    geneig_fixture()
    exact_fixture()
