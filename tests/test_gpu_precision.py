"""Reduced-precision forms of the mean-field contraction against the FP64 oracle.

Bars: the int8-sliced form keeps W within 1e-10 and the right-hand side within 1e-9 relative
(normwise, per quantity).  Split-FP32 and 3xBF16 are measured and their errors recorded;
their bounds below are the accuracy those schemes can reach with FP32 accumulation.
"""
import numpy as np
import pytest

from codyfortran_amd import hip
from codyfortran_amd import workloads as wl
from tests.common import oracle_system, rel, state_rel

pytestmark = pytest.mark.gpu

MODES = {
    "int8_sliced": (hip.PREC_INT8_SLICED, 1e-10, 1e-9),
    "split_fp32": (hip.PREC_SPLIT_FP32, 1e-6, 1e-5),
    "bf16x3": (hip.PREC_BF16X3, 1e-5, 1e-4),
}


@pytest.fixture(scope="module")
def full_size():
    w = wl.WORKLOADS["mctdhf_1d_4096"]
    s, x, wts = oracle_system(w)
    ctx, xg, wg = wl.make_context(w)
    state = wl.seeded_state(xg, wg, w.N, w.M, seed=1234)
    C, orbs = wl.split_state(state, w.nC, w.nG, w.M)
    ref_w = s.meanfield_w(orbs)
    ref_d = s.time_derivative(state)
    yield w, ctx, state, orbs, ref_w, ref_d, wts
    ctx.close()


@pytest.mark.parametrize("mode", list(MODES))
def test_contraction_and_rhs_error(full_size, mode):
    w, ctx, state, orbs, ref_w, ref_d, wts = full_size
    code, wtol, rtol = MODES[mode]
    ctx.set_precision(code)
    try:
        ew = rel(ctx.meanfield_w(orbs), ref_w)
        ed = state_rel(ctx.time_derivative(state), ref_d, w.nC, wts, w.M)
        pair_err = max(rel(ctx.meanfield_w(orbs)[:, p], ref_w[:, p]) for p in range(ref_w.shape[1]))
        print(f"\n{mode}: W rel err {ew:.3e} (worst pair {pair_err:.3e}), RHS rel err {ed:.3e}")
        assert ew < wtol and ed < rtol
    finally:
        ctx.set_precision(hip.PREC_FP64)
