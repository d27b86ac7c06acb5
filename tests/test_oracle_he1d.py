"""He1d known-answer pins for the oracle (the reference's MCTDHX simulation tests).

System of test/simulationTests/He1d/T_He1d_0[1-7]_*.json: uniform 21-point grid
on [-20, 20] (weights dx = 2), 9-point finite-difference kinetic energy with
zero ends, V = -2/(|x|+1), W = 1/(|x-x'|+1), two fermionic body types with one
particle each, unit masses, regularisation 1e-10 (the reference default).
"""
import numpy as np
import pytest

import oracle

E_ED = -2.4852443924542094        # T_He1d_01_ExactDiagGridBasedFullExpansion.f90:63
E_MCTDHX = -2.4852442075051160    # T_He1d_04_ImagTimePropagationMctdhx.f90:108
E_RT = -2.4852441374569345        # T_He1d_05_TimePropagationMctdhx.f90:122, T_He1d_06_...Sil.f90:135


def he1d(n_orbs):
    nG = 21
    x = -20.0 + 2.0 * np.arange(nG)
    w = np.full(nG, 2.0)
    s = oracle.System(x, w, [1, 1], [n_orbs, n_orbs], reg=1e-10)
    s.set_kinetic_fd(2.0)
    s.set_potential(-2.0 / (np.abs(x) + 1.0))
    s.set_interaction_softyukawa(1.0, 0.0, 1.0, 0.0)
    return s, x, w


def test_grid_exact_diagonalisation_independent_of_oracle():
    """Plain numpy two-particle grid Hamiltonian reproduces the reference's ED energy."""
    nG = 21
    x = -20.0 + 2.0 * np.arange(nG)
    c = [-205.0 / 72.0, 8.0 / 5.0, -1.0 / 5.0, 8.0 / 315.0, -1.0 / 560.0]
    D2 = sum(c[k] * (np.eye(nG, k=k) + (np.eye(nG, k=-k) if k else 0)) for k in range(5)) / 4.0
    h = -0.5 * D2 + np.diag(-2.0 / (np.abs(x) + 1.0))
    K = 1.0 / (np.abs(x[:, None] - x[None, :]) + 1.0)
    H = np.kron(h, np.eye(nG)) + np.kron(np.eye(nG), h) + np.diag(K.ravel())
    assert abs(np.linalg.eigvalsh(H)[0] - E_ED) < 1e-7


def test_complete_basis_ci_energy_matches_reference_ed():
    """With a complete orbital set the oracle's CI problem is the exact problem."""
    s, x, w = he1d(21)
    nG = 21
    basis = np.diag(1.0 / np.sqrt(w)).astype(np.complex128)
    orbs = np.concatenate([basis, basis], axis=1)
    h1 = s.fill_h1(orbs)
    h2 = s.fill_h2(orbs)
    H = np.zeros((s.nC, s.nC), dtype=np.complex128)
    for i in range(s.nC):
        e = np.zeros(s.nC, dtype=np.complex128)
        e[i] = 1.0
        H[:, i] = s.apply_h1(e, h1) + s.apply_h2(e, h2)
    ev, U = np.linalg.eigh(H)
    assert abs(ev[0] - E_ED) < 1e-7
    state = np.concatenate([U[:, 0], orbs.T.ravel()])
    assert abs(s.energy(state) - ev[0]) < 1e-12


def _ground_state(s, x, w, n_orbs):
    # start: lowest one-body eigenfunctions for both body types, C on the first configuration
    nG = len(x)
    h = np.zeros((nG, nG), dtype=np.complex128)
    for g in range(nG):
        e = np.zeros((nG, s.nO), dtype=np.complex128, order="F")
        e[g, 0] = 1.0 / np.sqrt(w[g])
        d, _ = s.apply_single_body(e)
        h[:, g] = d[:, 0] * np.sqrt(w[g])
    hs = np.sqrt(w)[:, None] * h / np.sqrt(w)[None, :]
    _, U = np.linalg.eigh(0.5 * (hs + hs.conj().T))
    phi = U[:, :n_orbs] / np.sqrt(w)[:, None]
    rng = np.random.default_rng(0)
    phi2 = phi + 0.01 * rng.normal(size=phi.shape)
    orbs = s.orthonormalize(np.concatenate([phi, phi2], axis=1))
    C = np.zeros(s.nC, dtype=np.complex128)
    C[0] = 1.0
    C[1:] = 1e-3
    C = s.normalize(C)
    state = np.concatenate([C, orbs.T.ravel()])
    # imaginary-time RK1 with renormalisation (T_He1d_04_ImagTimePropagationMctdhx.f90:82-99)
    dt, e_old = 0.05, 0.0
    for outer in range(4000):
        for _ in range(10):
            state = s.rk1(state, 0.0, dt, rhs=oracle.RHS_IMAG)
            C, o = s.split(state)
            C = s.normalize(C)
            o = s.orthonormalize(o)
            state = np.concatenate([C, o.T.ravel()])
        e = s.energy(state)
        if abs(e - e_old) < 1e-13:
            break
        e_old = e
    return state, e


@pytest.fixture(scope="module")
def he1d_ground():
    s, x, w = he1d(4)
    state, e = _ground_state(s, x, w, 4)
    return s, state, e


def test_imaginary_time_mctdhx_energy(he1d_ground):
    s, state, e = he1d_ground
    assert abs(e - E_MCTDHX) < 1e-7, e


def test_real_time_split_step_energy_and_norm(he1d_ground):
    """T_He1d_05/06: Strang split (RK4/RK4 and SIL/RK4), dt = 1e-5 to t = 1e-3."""
    s, state0, _ = he1d_ground
    n0 = np.linalg.norm(state0)
    e0 = s.energy(state0)
    st_sil = state0.copy()
    t = 0.0
    dt = 1e-5
    for _ in range(100):
        st_sil, ok = s.strang(st_sil, t, t + dt, krylov_dim=24, tol=1e-12, max_restarts=8)
        assert ok
        t += dt
    e_sil = s.energy(st_sil)
    assert abs(e_sil - E_RT) < 1e-7
    assert abs(e_sil - e0) < 1e-6
    assert abs(np.linalg.norm(st_sil) - n0) < 1e-10
