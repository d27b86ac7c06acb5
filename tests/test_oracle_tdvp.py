"""Pins for the oracle's full MCTDHX right-hand side from the Dirac-Frenkel principle.

For an MCTDHF state Psi(C, phi) the exact variational equations of motion
satisfy  i dPsi/dt = P_T (H - E) Psi,  with P_T the orthogonal projector (in
the grid metric) onto the tangent space of the MCTDHF manifold, and the
reference's gauge (orbital derivatives orthogonal to the orbitals, CI
derivative orthogonal to C).  The check builds Psi on the N-particle grid
explicitly (tests/fock.py), so a dropped term, a wrong sign, a lost factor
1/2 or a transposed density-matrix index in the oracle fails it.

The reference's code applies the complex conjugate of the CI Hamiltonian and
forms rho1, rho2 as <a+_i a_j>, <a+_i1 a+_i2 a_j2 a_j1> (see
tests/test_oracle_ci.py).  For real C and real orbitals this is immaterial;
for complex data the reference RHS equals the Dirac-Frenkel RHS of the state
with conjugated CI vector, with the CI derivative mapped dC -> -conj(dC).
Both forms are pinned here.
"""
import itertools

import numpy as np
import pytest

import oracle
from tests import fock


def _make(N, M, nE, nLoc, L, seed, complex_data, reg=1e-14):
    x, w = oracle.fedvr_grid(-L, L, nE, nLoc)
    s = oracle.System(x, w, [N], [M], reg=reg)
    s.set_kinetic_fedvr(-L, L, nE, nLoc)
    s.set_potential(-2.0 / np.sqrt(x ** 2 + 1.0))
    s.set_interaction_softyukawa(1.0, 1.0, 0.0, 0.0)
    rng = np.random.default_rng(seed)
    nG = len(x)
    orbs = rng.normal(size=(nG, M)) * np.exp(-x ** 2 / 8)[:, None]
    C = rng.normal(size=s.nC)
    if complex_data:
        orbs = orbs + 1j * rng.normal(size=(nG, M)) * np.exp(-x ** 2 / 8)[:, None]
        C = C + 1j * rng.normal(size=s.nC)
    orbs = s.orthonormalize(orbs.astype(np.complex128))
    C = (C / np.linalg.norm(C)).astype(np.complex128)
    return s, x, w, orbs, C


def _one_body_matrix(s):
    nG = s.nG
    h = np.zeros((nG, nG), dtype=np.complex128)
    for g in range(nG):
        e = np.zeros((nG, s.nO), dtype=np.complex128, order="F")
        e[g, 0] = 1.0
        d, _ = s.apply_single_body(e)
        h[:, g] = d[:, 0]
    return h


def _tangent_check(s, x, w, orbs, C, dC, dorbs, N):
    M = orbs.shape[1]
    nG = len(x)
    combs = list(itertools.combinations(range(M), N))
    h = _one_body_matrix(s)
    K = oracle.softyukawa(x[:, None] - x[None, :], 1.0, 1.0, 0.0, 0.0)
    psi = fock.many_body(orbs, C, N)
    D = np.ones((nG,) * N)
    for ax in range(N):
        shape = [1] * N
        shape[ax] = nG
        D = D * np.sqrt(w).reshape(shape)
    Hpsi = fock.apply_one_body(h, psi) + fock.apply_pair(K, psi)
    E = np.vdot(D * psi, D * Hpsi) / np.vdot(D * psi, D * psi)
    # tangent vectors: dPsi/dC_I and dPsi/dphi_k[e_g]
    tv = [fock.slater(orbs, c).ravel() for c in combs]
    for k in range(M):
        for g in range(nG):
            t = 0
            for cI, comb in zip(C, combs):
                if k not in comb:
                    continue
                o2 = orbs.copy()
                o2[:, k] = 0.0
                o2[g, k] = 1.0
                t = t + cI * fock.slater(o2, comb)
            if not np.isscalar(t):
                tv.append(t.ravel())
    B = np.array(tv).T * D.ravel()[:, None]
    U, sv, _ = np.linalg.svd(B, full_matrices=False)
    Q = U[:, sv > 1e-10 * sv[0]]
    target = D.ravel() * (Hpsi - E * psi).ravel()
    proj = Q @ (Q.conj().T @ target)
    # dPsi from the RHS (product rule)
    dpsi = sum(d * fock.slater(orbs, c) for d, c in zip(dC, combs))
    for k in range(M):
        for cI, comb in zip(C, combs):
            if k not in comb:
                continue
            o2 = orbs.copy()
            o2[:, k] = dorbs[:, k]
            dpsi = dpsi + cI * fock.slater(o2, comb)
    lhs = D.ravel() * (1j * dpsi).ravel()
    return np.linalg.norm(lhs - proj) / np.linalg.norm(proj)


@pytest.mark.parametrize("N,M,nE,nLoc,L", [(2, 4, 4, 5, 6.0), (3, 5, 3, 5, 5.0)])
def test_rhs_real_data_is_dirac_frenkel(N, M, nE, nLoc, L):
    s, x, w, orbs, C = _make(N, M, nE, nLoc, L, seed=10 + N, complex_data=False)
    state = np.concatenate([C, orbs.T.ravel()])
    d = s.time_derivative(state)
    dC, dorbs = s.split(d)
    err = _tangent_check(s, x, w, orbs, C, dC, dorbs, N)
    assert err < 1e-9, err


@pytest.mark.parametrize("N,M,nE,nLoc,L", [(2, 4, 4, 5, 6.0), (3, 5, 3, 5, 5.0)])
def test_rhs_complex_data_is_dirac_frenkel_of_conjugated_ci(N, M, nE, nLoc, L):
    s, x, w, orbs, C = _make(N, M, nE, nLoc, L, seed=20 + N, complex_data=True)
    state = np.concatenate([C, orbs.T.ravel()])
    d = s.time_derivative(state)
    dC, dorbs = s.split(d)
    err = _tangent_check(s, x, w, orbs, np.conj(C), -np.conj(dC), dorbs, N)
    assert err < 1e-9, err
    # and the textbook reading (no conjugation) does not hold for complex data
    err_plain = _tangent_check(s, x, w, orbs, C, dC, dorbs, N)
    assert err_plain > 1e-4


def test_rhs_gauge_and_norm_invariants():
    s, x, w, orbs, C = _make(3, 6, 5, 6, 8.0, seed=3, complex_data=True, reg=1e-8)
    state = np.concatenate([C, orbs.T.ravel()])
    d = s.time_derivative(state)
    dC, dorbs = s.split(d)
    assert abs(np.vdot(C, dC)) < 1e-14  # CI derivative orthogonal to C (norm conservation)
    ov = orbs.conj().T @ (dorbs * w[:, None])
    assert np.max(np.abs(ov)) < 1e-13  # gauge <phi_i|dphi_j>_w = 0
    # linear split part is -i (T+V) phi, unprojected, CI frozen
    dl = s.time_derivative_orbs_lin(state)
    assert np.all(dl[: s.nC] == 0)
    hphi, _ = s.apply_single_body(orbs)
    np.testing.assert_allclose(s.split(dl)[1], -1j * hphi, atol=1e-13)


def test_regularized_inverse_closed_form():
    """regInvMat = rho1 (rho1^2 + eps)^{-1} for Hermitian rho1 (gauge-invariant closed form)."""
    s, x, w, orbs, C = _make(3, 6, 5, 6, 8.0, seed=4, complex_data=True, reg=1e-6)
    r1 = s.rdm1(C)
    R = s.reg_inverse(r1)
    ref = r1 @ np.linalg.inv(r1 @ r1 + 1e-6 * np.eye(6))
    np.testing.assert_allclose(R, ref, atol=1e-10)
