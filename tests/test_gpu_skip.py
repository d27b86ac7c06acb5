"""Zero-source skipping of the mean-field contraction (cody_set_source_skip), both contraction forms.

The reference leaves zero sources out of its convolution sums
(src/Utils/M_Utils_ConvolutionIntegral.f90:84-93); the contraction kernels skip every k block
whose pair sources are all exactly zero.  The skipped products are exact zeros, so the mean fields
with and without skipping differ only by summation order, and both match the oracle.  Cases: the
seeded states (pair sources zero beyond |x| ~ 27 on the 401-point grid and ~ 55 on the 4096-point
grid), orbitals confined to a few points (one active block, padded lists), and zero orbitals (no
active block: W = 0 exactly).
"""
import numpy as np
import pytest

from codyfortran_amd import hip
from codyfortran_amd import workloads as wl
from tests.common import oracle_system, rel, state_rel

pytestmark = pytest.mark.gpu

FORMS = {"fp64": hip.PREC_FP64, "int8_sliced": hip.PREC_INT8_SLICED}
W_BAR = {"fp64": 1e-13, "int8_sliced": 1e-10}


@pytest.fixture(scope="module", params=["mctdhf_1d_401", "mctdhf_1d_4096"])
def system(request):
    w = wl.WORKLOADS[request.param]
    s, x, wts = oracle_system(w)
    ctx, xg, wg = wl.make_context(w)
    print(f"\n[skip: {request.param}]", flush=True)
    yield w, s, ctx, xg, wts
    ctx.close()


@pytest.mark.parametrize("form", list(FORMS))
def test_skip_matches_dense_and_oracle(system, form):
    w, s, ctx, x, wts = system
    state = wl.seeded_state(x, wts, w.N, w.M, seed=31)
    C, orbs = wl.split_state(state, w.nC, w.nG, w.M)
    ctx.set_precision(FORMS[form])
    try:
        ctx.set_source_skip(True)
        w_skip, d_skip = ctx.meanfield_w(orbs), ctx.time_derivative(state)
        ctx.set_source_skip(False)
        w_dense, d_dense = ctx.meanfield_w(orbs), ctx.time_derivative(state)
    finally:
        ctx.set_source_skip(True)
        ctx.set_precision(hip.PREC_FP64)
    w_orc, d_orc = s.meanfield_w(orbs), s.time_derivative(state)
    print(f"  {form}: skip vs dense W {rel(w_skip, w_dense):.2e}; W vs oracle {rel(w_skip, w_orc):.2e}")
    assert rel(w_skip, w_dense) < 1e-15 * max(1.0, 1e5 * W_BAR[form])  # summation order only
    assert rel(w_skip, w_orc) < W_BAR[form]
    assert state_rel(d_skip, d_orc, w.nC, wts, w.M) < 1e-12
    assert state_rel(d_dense, d_orc, w.nC, wts, w.M) < 1e-12


@pytest.mark.parametrize("form", list(FORMS))
def test_confined_and_zero_orbitals(system, form):
    w, s, ctx, x, wts = system
    rng = np.random.default_rng(8)
    ctx.set_precision(FORMS[form])
    try:
        # orbitals living on 5 points around the grid's middle: one or two active blocks
        mid = w.nG // 2
        orbs = np.zeros((w.nG, w.M), dtype=np.complex128, order="F")
        orbs[mid - 2: mid + 3, :] = rng.standard_normal((5, w.M)) + 1j * rng.standard_normal((5, w.M))
        w_orc = s.meanfield_w(orbs)
        got = ctx.meanfield_w(orbs)
        assert rel(got, w_orc) < W_BAR[form]
        # zero orbitals: no active block, the mean fields are exactly zero
        zero = np.zeros((w.nG, w.M), dtype=np.complex128, order="F")
        assert np.array_equal(ctx.meanfield_w(zero), np.zeros((w.nG, w.M * (w.M + 1) // 2), dtype=np.complex128))
        # and the context still gives the dense result afterwards
        ctx.set_source_skip(False)
        assert rel(ctx.meanfield_w(orbs), w_orc) < W_BAR[form]
    finally:
        ctx.set_source_skip(True)
        ctx.set_precision(hip.PREC_FP64)
