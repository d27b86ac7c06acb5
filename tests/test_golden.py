"""The committed fixture tests/golden/mctdhf_1d_401.npz (written by tests/golden/make_golden.py
from the FP64 oracle) against the oracle today (CPU) and against the HIP path through the C ABI
(GPU).  Bars: the oracle reproduces its own output to rounding (its OpenMP sums are in fixed
order); the HIP path to the FP64 parity bar, 1e-12 relative."""
import os

import numpy as np
import pytest

from codyfortran_amd import workloads as wl
from tests.common import oracle_system, rel, state_rel

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mctdhf_1d_401.npz")
W = wl.WORKLOADS["mctdhf_1d_401"]


@pytest.fixture(scope="module")
def golden():
    return np.load(FIX, allow_pickle=False)


def test_fixture_is_the_seeded_state(golden):
    x, w = golden["x"], golden["w"]
    np.testing.assert_array_equal(golden["state"], wl.seeded_state(x, w, W.N, W.M, seed=1234))


def test_oracle_reproduces_fixture(golden):
    s, x, wts = oracle_system(W)
    np.testing.assert_allclose(x, golden["x"], rtol=0, atol=1e-13)
    st = golden["state"]
    assert state_rel(s.time_derivative(st), golden["rhs"], W.nC, wts, W.M) < 1e-13
    assert state_rel(s.time_derivative_nonlin(st), golden["rhs_nonlin"], W.nC, wts, W.M) < 1e-13
    assert abs(s.energy(st) - float(golden["energy"])) < 1e-12
    assert abs(s.dipole(st) - float(golden["dipole"])) < 1e-12
    _, orbs = wl.split_state(st, W.nC, W.nG, W.M)
    assert rel(s.meanfield_w(orbs), golden["W"]) < 1e-13


@pytest.mark.gpu
def test_hip_path_matches_fixture(golden):
    ctx, x, wts = wl.make_context(W)
    try:
        st = golden["state"]
        assert state_rel(ctx.time_derivative(st), golden["rhs"], W.nC, wts, W.M) < 1e-12
        assert abs(ctx.energy(st) - float(golden["energy"])) < 1e-12 * abs(float(golden["energy"]))
        assert abs(ctx.dipole(st) - float(golden["dipole"])) < 1e-12
        _, orbs = wl.split_state(st, W.nC, W.nG, W.M)
        assert rel(ctx.meanfield_w(orbs), golden["W"]) < 1e-12
    finally:
        ctx.close()
