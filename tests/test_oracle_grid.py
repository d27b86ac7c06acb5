"""Pins for the oracle's grid pieces: GLL quadrature, FEDVR grid, FEDVR derivatives.

Known answers come from the reference's own component test
(test/componentTests/Utils/T_Utils_DerivativeFedvr.f90:20-126) and from the
mathematics (numpy's Legendre routines, polynomial exactness, self-adjointness).
"""
import numpy as np
import numpy.polynomial.legendre as L
import pytest

import oracle


@pytest.mark.parametrize("n", [2, 3, 5, 6, 9, 21])
def test_gll_matches_numpy_legendre(n):
    x, w = oracle.gll(n)
    # GLL interior nodes are the roots of P'_{n-1}
    c = np.zeros(n)
    c[-1] = 1.0
    inner = np.sort(L.legroots(L.legder(c))) if n > 2 else np.array([])
    ref = np.concatenate([[-1.0], inner, [1.0]])
    np.testing.assert_allclose(x, ref, atol=1e-14)
    Pn1 = L.legval(ref, c)
    np.testing.assert_allclose(w, 2.0 / (n * (n - 1) * Pn1 ** 2), rtol=1e-13)


@pytest.mark.parametrize("n", [3, 5, 6, 11])
def test_gll_polynomial_exactness(n):
    a, b = -0.7, 2.3
    x, w = oracle.gll(n, a, b)
    assert abs(w.sum() - (b - a)) < 1e-14
    for k in range(2 * n - 2):  # exact for degree <= 2n-3
        exact = (b ** (k + 1) - a ** (k + 1)) / (k + 1)
        assert abs(np.dot(w, x ** k) - exact) < 1e-12 * max(1.0, abs(exact))
    k = 2 * n - 2  # degree 2n-2 is not integrated exactly
    exact = (b ** (k + 1) - a ** (k + 1)) / (k + 1)
    assert abs(np.dot(w, x ** k) - exact) > 1e-10


def test_fedvr_grid_layout():
    # 819 elements x 6 GLL points with both ends kept: 819*5+1 = 4096 points
    x, w = oracle.fedvr_grid(-200.0, 200.0, 819, 6)
    assert len(x) == 4096
    assert np.all(np.diff(x) > 0)
    assert abs(w.sum() - 400.0) < 1e-10
    # excluded ends drop one point each (src/Utils/M_Utils_Fedvr.f90:108-111)
    x2, _ = oracle.fedvr_grid(0.0, 20.0, 10, 21, True, False)
    assert len(x2) == 10 * 20 and x2[0] > 0.0 and x2[-1] == 20.0


def test_fedvr_second_derivative_reference_kat():
    """T_Utils_DerivativeFedvr.f90:20-69: r e^{-2r} on [0,20], 10 el x 21 pts, xmin excluded, rel err < 1e-8."""
    xmin, xmax, nE, nLoc = 0.0, 20.0, 10, 21
    x, _ = oracle.fedvr_grid(xmin, xmax, nE, nLoc, True, False)
    lam = 2.0
    f = x * np.exp(-lam * x)
    exp2 = (-2.0 * lam + lam ** 2 * x) * np.exp(-lam * x)
    d2 = oracle.fedvr_derivative(2, f, xmin, xmax, nE, nLoc, True, False)
    rel = np.max(np.abs(d2 - exp2)) / np.max(np.abs(exp2))
    assert rel < 1e-8


def test_fedvr_first_derivative_reference_kat():
    """T_Utils_DerivativeFedvr.f90:77-126."""
    xmin, xmax, nE, nLoc = 0.0, 20.0, 10, 21
    x, _ = oracle.fedvr_grid(xmin, xmax, nE, nLoc, True, False)
    lam = 2.0
    f = x * np.exp(-lam * x)
    exp1 = (1.0 - lam * x) * np.exp(-lam * x)
    d1 = oracle.fedvr_derivative(1, f, xmin, xmax, nE, nLoc, True, False)
    assert np.max(np.abs(d1 - exp1)) / np.max(np.abs(exp1)) < 1e-8


def test_fedvr_kref_properties():
    K, F = oracle.fedvr_kref(6)
    np.testing.assert_allclose(K, K.T, atol=1e-13)
    # constants are in the kernel of the stiffness matrix; K is negative semi-definite
    np.testing.assert_allclose(K.sum(axis=1), 0.0, atol=1e-12)
    ev = np.linalg.eigvalsh(K)
    assert ev.max() < 1e-12 and (ev < -1e-8).sum() == 5


def test_fedvr_second_derivative_weighted_self_adjoint():
    """<f|D2 g>_w = conj(<g|D2 f>_w) to 1e-12 (pattern of the reference's FEDVR kinetic tests)."""
    xmin, xmax, nE, nLoc = -10.0, 10.0, 7, 6
    x, w = oracle.fedvr_grid(xmin, xmax, nE, nLoc)
    rng = np.random.default_rng(3)
    f = (rng.normal(size=len(x)) + 1j * rng.normal(size=len(x))) * np.exp(-x ** 2 / 8)
    g = (rng.normal(size=len(x)) + 1j * rng.normal(size=len(x))) * np.exp(-x ** 2 / 8)
    d2g = oracle.fedvr_derivative(2, g, xmin, xmax, nE, nLoc)
    d2f = oracle.fedvr_derivative(2, f, xmin, xmax, nE, nLoc)
    a = np.sum(np.conj(f) * d2g * w)
    b = np.conj(np.sum(np.conj(g) * d2f * w))
    assert abs(a - b) < 1e-12 * max(1.0, abs(a))


def test_symeig_matches_numpy():
    rng = np.random.default_rng(0)
    for n in (1, 2, 5, 16, 24):
        a = rng.normal(size=(n, n))
        a = a + a.T
        ev, V = oracle.symeig(a)
        np.testing.assert_allclose(np.sort(ev), np.linalg.eigvalsh(a), atol=1e-12)
        np.testing.assert_allclose(V @ np.diag(ev) @ V.T, a, atol=1e-12)
