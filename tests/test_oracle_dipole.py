"""Pins the oracle's dipole <x> (PrinterObservableLinear_DumpDipole,
src/Utils/M_Utils_PrinterObservableLinear.f90:25-70: sum_ij rho1(i,j) <phi_j|x phi_i>_w) against
the expectation value of X = sum_n x_n over the explicit N-particle grid wave function
(tests/fock.py), and against the closed form of a single determinant."""
import itertools

import numpy as np
import pytest

import oracle
from tests import fock
from tests.test_oracle_tdvp import _make


def _brute_dipole(x, w, orbs, C, N):
    psi = fock.many_body(orbs, C, N)
    nG = len(x)
    D = np.ones((nG,) * N)
    X = np.zeros((nG,) * N)
    for ax in range(N):
        shape = [1] * N
        shape[ax] = nG
        D = D * w.reshape(shape)
        X = X + x.reshape(shape)
    return float(np.real(np.sum(D * X * np.abs(psi) ** 2) / np.sum(D * np.abs(psi) ** 2)))


@pytest.mark.parametrize("N,M,nE,nLoc,L", [(2, 4, 4, 5, 6.0), (3, 5, 3, 5, 5.0)])
def test_dipole_real_data_is_the_many_body_expectation(N, M, nE, nLoc, L):
    s, x, w, orbs, C = _make(N, M, nE, nLoc, L, seed=30 + N, complex_data=False)
    orbs = orbs * np.exp(0.3 * x)[:, None]  # break the parity of the envelope: <x> != 0
    orbs = s.orthonormalize(orbs.astype(np.complex128))
    state = np.concatenate([C, orbs.T.ravel()])
    d = s.dipole(state)
    assert abs(d) > 1e-2
    assert abs(d - _brute_dipole(x, w, orbs, C, N)) < 1e-12


@pytest.mark.parametrize("N,M,nE,nLoc,L", [(2, 4, 4, 5, 6.0), (3, 5, 3, 5, 5.0)])
def test_dipole_complex_data(N, M, nE, nLoc, L):
    s, x, w, orbs, C = _make(N, M, nE, nLoc, L, seed=40 + N, complex_data=True)
    orbs = s.orthonormalize((orbs * np.exp(0.3 * x)[:, None]).astype(np.complex128))
    state = np.concatenate([C, orbs.T.ravel()])
    d = s.dipole(state)
    # the reference's rho1 is <a+_i a_j> (the conjugate of its docstring's convention), so for
    # complex data its <x> is that of the state with the conjugated CI vector
    assert abs(d - _brute_dipole(x, w, orbs, np.conj(C), N)) < 1e-12
    assert abs(d - _brute_dipole(x, w, orbs, C, N)) > 1e-3


def test_dipole_of_one_determinant_is_the_sum_of_orbital_centres():
    x, w = oracle.fedvr_grid(-14.0, 14.0, 14, 6)
    s = oracle.System(x, w, [2], [3])
    centres = [-8.0, 0.5, 8.0]  # far apart: Gram-Schmidt changes them by < 1e-12
    orbs = np.stack([np.exp(-(x - c) ** 2) * np.exp(0.4j * x) for c in centres], axis=1)
    orbs = s.orthonormalize(orbs.astype(np.complex128))
    # occupy orbitals 0 and 2: a single determinant, so <x> = <phi_0|x|phi_0> + <phi_2|x|phi_2>
    combs = list(itertools.combinations(range(3), 2))
    C = np.zeros(len(combs), dtype=np.complex128)
    C[combs.index((0, 2))] = np.exp(0.3j)
    state = np.concatenate([C, orbs.T.ravel()])
    ref = sum(np.sum(np.abs(orbs[:, k]) ** 2 * x * w) for k in (0, 2))
    assert abs(s.dipole(state) - ref) < 1e-12
    assert abs(s.dipole(state) - (centres[0] + centres[2])) < 1e-8
