"""Brute-force second quantisation and first-quantised many-body algebra.

Independent of the oracle's configuration tables: fermionic operators are
dense Jordan-Wigner matrices on the full 2^M Fock space, and many-body wave
functions are explicit Slater-determinant expansions on the grid.  Used by
the tests to pin the oracle's conventions (reduced density matrices, CI
H-apply) and its equations of motion (Dirac-Frenkel residual).
"""
from __future__ import annotations

import itertools
import math

import numpy as np


def annihilators(n_modes: int):
    """a_p as dense 2^n x 2^n matrices; basis index = occupation bit pattern (bit p = mode p)."""
    dim = 1 << n_modes
    ops = []
    for p in range(n_modes):
        a = np.zeros((dim, dim))
        for occ in range(dim):
            if occ >> p & 1:
                sign = -1.0 if bin(occ & ((1 << p) - 1)).count("1") % 2 else 1.0
                a[occ ^ (1 << p), occ] = sign
        ops.append(a)
    return ops


def lex_codes(n_orbs: int, n_bodies: int):
    """Occupation bit patterns of the N-of-M configurations in lexicographic order."""
    return [sum(1 << i for i in comb) for comb in itertools.combinations(range(n_orbs), n_bodies)]


def embed(coeffs, codes_per_type, shifts, n_modes):
    """CI vector (body type 0 fastest) -> Fock-space vector over all modes."""
    dim = 1 << n_modes
    v = np.zeros(dim, dtype=np.complex128)
    sizes = [len(c) for c in codes_per_type]
    for idx, c in enumerate(coeffs):
        code = 0
        rem = idx
        for codes, sh, sz in zip(codes_per_type, shifts, sizes):
            code |= codes[rem % sz] << sh
            rem //= sz
        v[code] = c
    return v


def restrict(vec, codes_per_type, shifts):
    sizes = [len(c) for c in codes_per_type]
    n = int(np.prod(sizes))
    out = np.zeros(n, dtype=np.complex128)
    for idx in range(n):
        code = 0
        rem = idx
        for codes, sh, sz in zip(codes_per_type, shifts, sizes):
            code |= codes[rem % sz] << sh
            rem //= sz
        out[idx] = vec[code]
    return out


# ---------------------------------------------------------------------------
# first quantisation on a grid (single body type of fermions)
# ---------------------------------------------------------------------------

def slater(orbs_cols, occupied):
    """Normalised Slater determinant on the product grid: (1/sqrt(N!)) det[phi_{o_a}(x_b)]."""
    N = len(occupied)
    nG = orbs_cols.shape[0]
    psi = np.zeros((nG,) * N, dtype=np.complex128)
    for perm in itertools.permutations(range(N)):
        sign = np.linalg.det(np.eye(N)[list(perm)])
        term = np.ones((1,) * 0, dtype=np.complex128)
        term = None
        for b, a in enumerate(perm):
            f = orbs_cols[:, occupied[a]]
            shape = [1] * N
            shape[b] = nG
            f = f.reshape(shape)
            term = f if term is None else term * f
        psi = psi + sign * term
    return psi / math.sqrt(math.factorial(N))


def many_body(orbs_cols, coeffs, n_bodies):
    n_orbs = orbs_cols.shape[1]
    psi = 0
    for c, comb in zip(coeffs, itertools.combinations(range(n_orbs), n_bodies)):
        psi = psi + c * slater(orbs_cols, comb)
    return psi


def apply_one_body(h, psi):
    """sum_i h acting on coordinate i of psi (h is nG x nG)."""
    N = psi.ndim
    out = np.zeros_like(psi)
    for ax in range(N):
        out = out + np.moveaxis(np.tensordot(h, psi, axes=([1], [ax])), 0, ax)
    return out


def apply_pair(K, psi):
    """sum_{i<j} K(x_i, x_j) psi (K is nG x nG, symmetric)."""
    N = psi.ndim
    nG = psi.shape[0]
    out = np.zeros_like(psi)
    for i in range(N):
        for j in range(i + 1, N):
            shape = [1] * N
            shape[i] = nG
            shape[j] = nG
            kk = K if i < j else K.T
            out = out + kk.reshape(shape) * psi
    return out
