"""The two Strang splits of the oracle (S_Propagator_SplitStep_Order2.f90:52-77).

A is the SIL on the linear part, -i (T + V + E x) phi on the orbitals
(TimeDerivativeOrbsLin, S_Method_Mb_OrbBased_Mctdhx.f90:152-177).
- The reference's B is TimeDerivativeCoeffsPlusOrbsNonLin (:180-226). It evolves C with the full
  h1 but leaves out -P h phi, so A + B is not the full right-hand side.
- The consistent B is the full right-hand side minus A.

Checks fixed by the mathematics: with the consistent B the step is second order, so its energy
error shrinks by 4 per halving of dt. The full RK4 conserves the energy to the regularisation
level. The reference split converges to a different flow: its energy error does not depend on dt.
"""
import numpy as np
import pytest

import oracle
from codyfortran_amd import workloads as wl
from tests.common import oracle_system, state_rel

W = wl.WORKLOADS["mctdhf_1d_401"]


@pytest.fixture(scope="module")
def system():
    s, x, wts = oracle_system(W)
    st = wl.smooth_state(x, wts, W.N, W.M, seed=1234)
    return s, wts, st


def test_consistent_nonlinear_part_is_full_minus_linear(system):
    s, wts, st = system
    d = s._rhs(s.lib.orc_time_derivative, st, 0.3) - s.time_derivative_orbs_lin(st, 0.3)
    # one RK1 step of length h on B is st + h B(st)
    h = 1e-3
    got = (s.rk1(st, 0.3, 0.3 + h, rhs=oracle.RHS_NONLIN_CONSISTENT) - st) / h
    assert state_rel(got, d, W.nC, wts, W.M) < 1e-12
    # and it differs from the reference's nonlinear part by -i P h phi
    ref_b = s.time_derivative_nonlin(st, 0.3)
    assert state_rel(ref_b, d, W.nC, wts, W.M) > 1e-3


def _energy_error(s, st, dt, consistent, T=0.5):
    metric = s.metric_weights()
    e0 = s.energy(st)
    y, t = st.copy(), 0.0
    for _ in range(int(round(T / dt))):
        y, ok = s.strang(y, t, t + dt, krylov_dim=W.krylov_dim, tol=1e-12, max_restarts=8, metric=metric,
                         consistent=consistent)
        assert ok
        t += dt
    return abs(s.energy(y) - e0) / abs(e0)


def test_consistent_split_is_second_order(system):
    s, wts, st = system
    e1 = _energy_error(s, st, 0.05, True)
    e2 = _energy_error(s, st, 0.025, True)
    assert e1 < 1e-4
    assert 3.5 < e1 / e2 < 4.5, (e1, e2)


def test_reference_split_error_does_not_shrink_with_the_step(system):
    s, wts, st = system
    e1 = _energy_error(s, st, 0.05, False)
    e2 = _energy_error(s, st, 0.025, False)
    assert e1 > 1e-3 and abs(e1 - e2) < 1e-2 * e1, (e1, e2)
