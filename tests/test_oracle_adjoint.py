"""The oracle's adjoint (vector-Jacobian products of the right-hand sides, the dipole gradient and
the discrete adjoint of the Chebyshev-Strang propagation) against central finite differences of
the oracle's own forward maps, and the Chebyshev half-step against the Lanczos one.

The reference has no adjoint: what pins these transposes is that every directional derivative
of the forward map (computed by the oracle's forward functions, which the other oracle tests pin
to the reference's conventions) equals the inner product with the computed gradient, for random
directions and cotangents.  A dropped term, a wrong conjugate or a transposed index changes that
inner product at O(1) and fails the test.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle

XMIN, XMAX, NE, NLOC = -8.0, 8.0, 10, 5


def _system(N=2, M=4, reg=1e-3, laser=True, T=2.0):
    x, w = oracle.fedvr_grid(XMIN, XMAX, NE, NLOC)
    s = oracle.System(x, w, [N], [M], reg=reg)
    s.set_kinetic_fedvr(XMIN, XMAX, NE, NLOC)
    s.set_potential(-2.0 / np.sqrt(x ** 2 + 1.0))
    s.set_interaction_softyukawa(1.0, 1.0, 0.0, 0.0)
    if laser:
        s.set_laser(np.array([0.3, -0.2, 0.15, 0.1]), T)
    return s, x, w


def _state(s, x, w, seed=3):
    rng = np.random.default_rng(seed)
    orbs = np.asfortranarray((rng.normal(size=(s.nG, s.nO)) + 1j * rng.normal(size=(s.nG, s.nO)))
                             * np.exp(-x[:, None] ** 2 / 8))
    orbs = s.orthonormalize(orbs)
    c = rng.normal(size=s.nC) + 1j * rng.normal(size=s.nC)
    c /= np.linalg.norm(c)
    return np.concatenate([c, orbs.T.reshape(-1)])


def _ip(s, a, b):
    """Re <a, b> in the metric of the adjoint (unit on C, grid weights on the orbitals)."""
    return float(np.real(np.sum(np.conj(a) * b * s.metric_weights())))


def _rand(s, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=s.n_state) + 1j * rng.normal(size=s.n_state)


RHS = {
    "full": (oracle.RHS_FULL, lambda s, y, t: s.time_derivative(y, t)),
    "lin": (oracle.RHS_LIN, lambda s, y, t: s.time_derivative_orbs_lin(y, t)),
    "nonlin": (oracle.RHS_NONLIN, lambda s, y, t: s.time_derivative_nonlin(y, t)),
    "consistent": (oracle.RHS_NONLIN_CONSISTENT,
                   lambda s, y, t: s.time_derivative(y, t) - s.time_derivative_orbs_lin(y, t)),
}


@pytest.mark.parametrize("part", list(RHS))
def test_rhs_vjp_matches_directional_derivative(part):
    code, f = RHS[part]
    s, x, w = _system()
    y = _state(s, x, w)
    t = 0.7
    h = 1e-6
    for seed in range(3):
        lam, dy = _rand(s, 10 + seed), _rand(s, 20 + seed)
        g, _ = s.rhs_vjp(y, lam, t, rhs=code)
        fd = (_ip(s, lam, f(s, y + h * dy, t)) - _ip(s, lam, f(s, y - h * dy, t))) / (2 * h)
        an = _ip(s, g, dy)
        assert abs(an - fd) <= 1e-7 * max(1.0, abs(fd)), (part, seed, an, fd)


@pytest.mark.parametrize("part", list(RHS))
def test_rhs_vjp_field_derivative(part):
    """<lam, df/dE> against a finite difference in one laser coefficient (dE/dc_1 = basis(t))."""
    code, f = RHS[part]
    s, x, w = _system()
    y = _state(s, x, w)
    t, T = 0.7, 2.0
    c0 = np.array([0.3, -0.2, 0.15, 0.1])
    lam = _rand(s, 5)
    _, eb = s.rhs_vjp(y, lam, t, rhs=code)
    h = 1e-5
    vals = []
    for sg in (1, -1):
        c = c0.copy()
        c[0] += sg * h
        s.set_laser(c, T)
        vals.append(_ip(s, lam, f(s, y, t)))
    s.set_laser(c0, T)
    basis = np.sin(np.pi * t / T) * np.sin(np.pi * t / T) ** 2
    fd = (vals[0] - vals[1]) / (2 * h) / basis
    assert abs(eb - fd) <= 1e-7 * max(1.0, abs(fd)), (part, eb, fd)
    if part in ("full", "lin", "consistent"):
        assert abs(eb) > 1e-3  # the field enters through h phi


def test_rhs_vjp_regularisation_term():
    """Small occupation numbers make the rho1 (rho1^2 + eps)^-1 derivative large: still exact."""
    s, x, w = _system(N=2, M=6, reg=1e-4)
    rng = np.random.default_rng(7)
    y = _state(s, x, w)
    c = np.zeros(s.nC, dtype=complex)  # nearly a single determinant: small natural occupations
    c[0] = 1.0
    c[1:] = 0.05 * (rng.normal(size=s.nC - 1) + 1j * rng.normal(size=s.nC - 1))
    y[: s.nC] = c / np.linalg.norm(c)
    lam, dy = _rand(s, 1), _rand(s, 2)
    g, _ = s.rhs_vjp(y, lam, 0.3)
    h = 1e-7
    fd = (_ip(s, lam, s.time_derivative(y + h * dy, 0.3)) - _ip(s, lam, s.time_derivative(y - h * dy, 0.3))) / (2 * h)
    assert abs(_ip(s, g, dy) - fd) <= 1e-6 * max(1.0, abs(fd))


def test_dipole_grad():
    s, x, w = _system()
    y = _state(s, x, w)
    g = s.dipole_grad(y)
    h = 1e-6
    for seed in range(3):
        dy = _rand(s, 30 + seed)
        fd = (s.dipole(y + h * dy) - s.dipole(y - h * dy)) / (2 * h)
        assert abs(_ip(s, g, dy) - fd) <= 1e-8 * max(1.0, abs(fd))


def _bounds(s):
    """Spectrum of h = T + V + E x for |E| <= sum |c_k| from the dense operator, widened by 5 %."""
    H = np.zeros((s.nG, s.nG), dtype=complex)
    for j in range(s.nG):
        e = np.zeros((s.nG, s.nO), dtype=complex, order="F")
        e[j, 0] = 1.0
        H[:, j] = s.apply_single_body(e, t=-1.0)[0][:, 0]  # t < 0: no field
    sw = np.sqrt(s.w)
    ev = np.linalg.eigvalsh((sw[:, None] * H / sw[None, :]).real)
    emax = 0.75
    lo = ev[0] - emax * np.max(np.abs(s.x))
    hi = ev[-1] + emax * np.max(np.abs(s.x))
    pad = 0.05 * (hi - lo)
    return lo - pad, hi + pad


@pytest.mark.parametrize("consistent", [False, True])
def test_chebyshev_strang_matches_lanczos_strang(consistent):
    s, x, w = _system()
    y = _state(s, x, w)
    lo, hi = _bounds(s)
    dt = 0.02
    K = s.cheb_degree(dt / 2, lo, hi)
    yc = s.strang_cheb(y, 0.4, 0.4 + dt, lo, hi, K, consistent=consistent)
    ys, ok = s.strang(y, 0.4, 0.4 + dt, krylov_dim=40, tol=1e-14, max_restarts=12, metric=s.metric_weights(),
                      consistent=consistent)
    d = yc - ys
    assert np.sqrt(_ip(s, d, d) / _ip(s, ys, ys)) < 1e-11


def test_cheb_degree_grows_with_tau():
    s, _, _ = _system()
    k1 = s.cheb_degree(0.01, -10.0, 1000.0)
    k2 = s.cheb_degree(0.02, -10.0, 1000.0)
    assert 5 < k1 < k2 < 60


@pytest.mark.parametrize("consistent", [False, True])
def test_dipole_gradient_matches_finite_differences(consistent):
    s, x, w = _system(T=0.6)
    y = _state(s, x, w)
    lo, hi = _bounds(s)
    dt, n = 0.05, 12
    K = s.cheb_degree(dt / 2, lo, hi)
    c0 = np.array([0.3, -0.2, 0.15, 0.1])
    s.set_laser(c0, 0.6)
    J, g, _ = s.dipole_gradient(y, 0.0, dt, n, lo, hi, K, consistent=consistent)
    h = 1e-6
    fd = np.zeros(len(c0))
    for k in range(len(c0)):
        vals = []
        for sg in (1, -1):
            c = c0.copy()
            c[k] += sg * h
            s.set_laser(c, 0.6)
            vals.append(s.dipole_gradient(y, 0.0, dt, n, lo, hi, K, consistent=consistent)[0])
        fd[k] = (vals[0] - vals[1]) / (2 * h)
    s.set_laser(c0, 0.6)
    assert np.linalg.norm(g - fd) <= 1e-6 * np.linalg.norm(fd), (g, fd)
    assert np.linalg.norm(fd) > 1e-3
