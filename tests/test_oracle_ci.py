"""Pins for the oracle's CI algebra against brute-force Jordan-Wigner matrices.

The reference's conventions, as its code implements them
(src/Coeffs/Generic/S_Coeffs_Generic.f90:103-542 with the excitation rule of
src/ConfigList/AllActive/Fermions/S_ConfigList_AllActive_Fermions.f90:232-330):
  rdm1(i,j)        = <Psi| a+_i a_j |Psi>
  rdm2(i1,i2,j1,j2) = <Psi| a+_i1 a+_i2 a_j2 a_j1 |Psi>
  H1-apply         = sum_{pq} h1(q,p) a+_p a_q
  H2-apply         = sum_{all} h2(i1,i2,j1,j2) a+_j1 a+_j2 a_i2 a_i1
Configurations are the lexicographic N-of-M combinations, body type 0 fastest.
"""
import numpy as np
import pytest

import oracle
from tests import fock


def _system(n_bodies, n_orbs):
    x = np.linspace(-1, 1, 3)
    return oracle.System(x, np.ones(3), n_bodies, n_orbs)


def _setup(n_bodies, n_orbs, seed):
    s = _system(n_bodies, n_orbs)
    nO = sum(n_orbs)
    shifts = np.cumsum([0] + list(n_orbs[:-1]))
    codes = [fock.lex_codes(m, n) for n, m in zip(n_bodies, n_orbs)]
    for bt in range(len(n_orbs)):
        assert list(s.config_codes(bt)) == codes[bt]
    rng = np.random.default_rng(seed)
    C = rng.normal(size=s.nC) + 1j * rng.normal(size=s.nC)
    C /= np.linalg.norm(C)
    a = fock.annihilators(nO)
    psi = fock.embed(C, codes, shifts, nO)
    return s, C, a, psi, codes, shifts, nO


CASES = [((2,), (4,)), ((3,), (5,)), ((1, 1), (4, 4)), ((2, 1), (4, 3))]


@pytest.mark.parametrize("nb,no", CASES)
def test_rdm1_rdm2_match_jordan_wigner(nb, no):
    s, C, a, psi, codes, shifts, nO = _setup(nb, no, 1)
    r1 = s.rdm1(C)
    r2 = s.rdm2(C)
    ad = [m.T for m in a]
    bt = np.repeat(np.arange(len(no)), no)
    for i in range(nO):
        for j in range(nO):
            ref = np.vdot(psi, ad[i] @ a[j] @ psi) if bt[i] == bt[j] else 0.0
            assert abs(r1[i, j] - ref) < 1e-13
    # <Psi| a+_i1 a+_i2 a_j2 a_j1 |Psi> = <a_i2 a_i1 Psi | a_j2 a_j1 Psi>, from the vectors
    # a_j2 a_j1 Psi (matrix-vector products only)
    aa = [[a[j2] @ (a[j1] @ psi) for j2 in range(nO)] for j1 in range(nO)]
    for i1 in range(nO):
        for i2 in range(nO):
            for j1 in range(nO):
                for j2 in range(nO):
                    ok = (bt[i1] == bt[j1] and bt[i2] == bt[j2]) or (bt[i1] == bt[j2] and bt[i2] == bt[j1])
                    ref = np.vdot(aa[i1][i2], aa[j1][j2]) if ok else 0.0
                    assert abs(r2[i1, i2, j1, j2] - ref) < 1e-13, (i1, i2, j1, j2)
    # invariants: Tr rho1 = N, sum rho2_{ijij} = N(N-1)   (src/Utils/M_Utils_PrinterObservable.f90:52-62)
    N = sum(nb)
    assert abs(np.trace(r1) - N) < 1e-13
    assert abs(np.einsum("ijij->", r2) - N * (N - 1)) < 1e-12


@pytest.mark.parametrize("nb,no", CASES)
def test_h_apply_matches_jordan_wigner(nb, no):
    s, C, a, psi, codes, shifts, nO = _setup(nb, no, 2)
    rng = np.random.default_rng(7)
    bt = np.repeat(np.arange(len(no)), no)
    same1 = bt[:, None] == bt[None, :]
    h1 = (rng.normal(size=(nO, nO)) + 1j * rng.normal(size=(nO, nO))) * same1
    h1 = h1 + h1.conj().T
    h2 = rng.normal(size=(nO,) * 4) + 1j * rng.normal(size=(nO,) * 4)
    # h2 blocks the reference fills: first-particle pair (i1,j1) and second (i2,j2) each within a body type
    mask = same1[:, None, :, None] & same1[None, :, None, :]
    h2 = h2 * mask
    ad = [m.T for m in a]
    # H1 Psi and H2 Psi by matrix-vector products: a_q Psi, a_i2 a_i1 Psi once, then the creators
    aq = [a[q] @ psi for q in range(nO)]
    h1psi = sum(h1[q, p] * (ad[p] @ aq[q]) for p in range(nO) for q in range(nO))
    aa = [[a[i2] @ aq[i1] for i2 in range(nO)] for i1 in range(nO)]
    h2psi = np.zeros_like(psi)
    for j1 in range(nO):
        for j2 in range(nO):
            v = sum(h2[i1, i2, j1, j2] * aa[i1][i2] for i1 in range(nO) for i2 in range(nO))
            h2psi = h2psi + ad[j1] @ (ad[j2] @ v)
    ref1 = fock.restrict(h1psi, codes, shifts)
    ref2 = fock.restrict(h2psi, codes, shifts)
    np.testing.assert_allclose(s.apply_h1(C, h1), ref1, atol=1e-12)
    np.testing.assert_allclose(s.apply_h2(C, h2), ref2, atol=1e-11)


def test_h_apply_is_conjugate_of_textbook_operator():
    """Documents the reference convention: for Hermitian h1 the CI operator is conj(textbook)."""
    s, C, a, psi, codes, shifts, nO = _setup((2,), (4,), 5)
    rng = np.random.default_rng(11)
    h1 = rng.normal(size=(nO, nO)) + 1j * rng.normal(size=(nO, nO))
    h1 = h1 + h1.conj().T
    ad = [m.T for m in a]
    textbook = sum(h1[p, q] * ad[p] @ a[q] for p in range(nO) for q in range(nO))
    got = s.apply_h1(C, h1)
    np.testing.assert_allclose(got, fock.restrict(textbook.conj() @ psi, codes, shifts), atol=1e-12)
    assert np.max(np.abs(got - fock.restrict(textbook @ psi, codes, shifts))) > 1e-3
