"""Reduced-precision forms of the mean-field contraction against the FP64 oracle (4096-point
grid, M = 8, N = 4, seeded state).

Every form is measured per quantity: the mean fields W, the full right-hand side and the energy,
each relative to the FP64 oracle.

- int8-sliced (five balanced base-254 digits, exact int8-MFMA products, FP64 recombination)
  meets the 1e-9 bar on all three quantities; W within 1e-10.
- BF16-sliced (the same digits as BF16 operands of BF16 MFMA, FP32 accumulation exact within each
  64-point block, FP64 recombination) meets the same bar, and its W and right-hand side equal the
  int8-sliced form's bit for bit.
- Split-FP32 and 3xBF16 accumulate their products in FP32 inside the MFMA, which rounds every
  partial sum at 2^-24 ~ 6e-8; their W errors sit there (measured 6.6e-8 and 5.7e-8).  They are
  kept as named regression bounds at those errors, and their 1e-9 test is skipped with that
  reason: a split-FP32 form that met 1e-9 would need exact products and chunk sums (pieces of
  at most 10 bits, six products per element) and could not beat the FP64 MFMA form, whose peak is
  half the FP32 one.
"""
import numpy as np
import pytest

from codyfortran_amd import hip
from codyfortran_amd import workloads as wl
from tests.common import oracle_system, rel, small, state_rel

pytestmark = pytest.mark.gpu

BAR = 1e-9  # every quantity of the forms the benchmark may report
# (W, RHS, energy) bounds per form
INT8_BOUNDS = (1e-10, BAR, BAR)
# regression bounds of the FP32-accumulating forms, at their measured W errors (6.6e-8, 5.7e-8)
SPLIT_FP32_REGRESSION = (1e-7, BAR, BAR)
BF16X3_REGRESSION = (1e-7, BAR, BAR)
MODES = {
    "int8_sliced": (hip.PREC_INT8_SLICED, INT8_BOUNDS),
    "bf16_sliced": (hip.PREC_BF16_SLICED, INT8_BOUNDS),
    "split_fp32": (hip.PREC_SPLIT_FP32, SPLIT_FP32_REGRESSION),
    "bf16x3": (hip.PREC_BF16X3, BF16X3_REGRESSION),
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
    ref_e = s.energy(state)
    yield w, ctx, state, orbs, ref_w, ref_d, ref_e, wts
    ctx.close()


def _errors(full_size, code):
    w, ctx, state, orbs, ref_w, ref_d, ref_e, wts = full_size
    ctx.set_precision(code)
    try:
        Wg = ctx.meanfield_w(orbs)
        ew = rel(Wg, ref_w)
        pair_err = max(rel(Wg[:, p], ref_w[:, p]) for p in range(ref_w.shape[1]))
        ed = state_rel(ctx.time_derivative(state), ref_d, w.nC, wts, w.M)
        ee = abs(ctx.energy(state) - ref_e) / abs(ref_e)
    finally:
        ctx.set_precision(hip.PREC_FP64)
    return ew, pair_err, ed, ee


@pytest.mark.parametrize("mode", list(MODES))
def test_contraction_rhs_and_energy_error(full_size, mode):
    code, (wb, rb, eb) = MODES[mode]
    ew, pair_err, ed, ee = _errors(full_size, code)
    print(f"\n{mode}: W rel err {ew:.3e} (worst pair {pair_err:.3e}), RHS rel err {ed:.3e}, energy rel err {ee:.3e}")
    assert ew < wb, ew
    assert ed < rb, ed
    assert ee < eb, ee


@pytest.mark.skip(reason="split-FP32 and 3xBF16 accumulate in FP32 inside the MFMA (rounding 2^-24 ~ 6e-8 per "
                         "partial sum), so W stays near 6e-8; the int8-sliced form is the one that meets 1e-9")
@pytest.mark.parametrize("mode", ["split_fp32", "bf16x3"])
def test_fp32_accumulating_forms_meet_the_bar(full_size, mode):
    code, _ = MODES[mode]
    ew, _, ed, ee = _errors(full_size, code)
    assert max(ew, ed, ee) < BAR


def test_bf16_sliced_equals_int8_sliced_bit_for_bit(full_size):
    """Every digit product and every block sum of a diagonal is an integer below 2^24, exact in the
    BF16 MFMA's FP32 accumulator as in the int8 MFMA's int32 one, and both forms recombine the
    blocks in the same FP64 order."""
    w, ctx, state, orbs, *_ = full_size
    out = {}
    for code in (hip.PREC_INT8_SLICED, hip.PREC_BF16_SLICED):
        ctx.set_precision(code)
        out[code] = (ctx.meanfield_w(orbs), ctx.time_derivative(state), ctx.energy(state))
    ctx.set_precision(hip.PREC_FP64)
    a, b = out[hip.PREC_INT8_SLICED], out[hip.PREC_BF16_SLICED]
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[1], b[1])
    assert a[2] == b[2]


def test_bf16_sliced_refuses_more_than_80_pair_columns():
    """12 orbitals (160 pair columns) are outside the BF16-sliced kernel: refused on the host."""
    w = small("m12_refuse", -10.0, 10.0, 32, 5, 6, 12)
    ctx, x, wts = wl.make_context(w)
    try:
        with pytest.raises(hip.CodyError, match="80 pair"):
            ctx.set_precision(hip.PREC_BF16_SLICED)
        ctx.set_precision(hip.PREC_INT8_SLICED)  # (the int8 form takes them, without the epilogue)
    finally:
        ctx.close()


@pytest.mark.parametrize("M", [10, 12])
def test_int8_above_eight_orbitals_matches_oracle(M):
    """Above 8 orbitals (up to 160 pair columns) the int8 form runs the 16-row contraction without
    the correlation epilogue, then the FP64 form's correlation on FP64 MFMA."""
    w = small(f"m{M}_int8", -10.0, 10.0, 32, 5, M // 2, M)
    s, x, wts = oracle_system(w)
    ctx, xg, wg = wl.make_context(w)
    state = wl.seeded_state(xg, wg, w.N, w.M, seed=77)
    _, orbs = wl.split_state(state, w.nC, w.nG, w.M)
    try:
        ctx.set_precision(hip.PREC_INT8_SLICED)
        ew = rel(ctx.meanfield_w(orbs), s.meanfield_w(orbs))
        ed = state_rel(ctx.time_derivative(state), s.time_derivative(state), w.nC, wts, w.M)
        ee = abs(ctx.energy(state) - s.energy(state)) / abs(s.energy(state))
    finally:
        ctx.close()
    print(f"\nint8, M = {M}: W rel err {ew:.3e}, RHS rel err {ed:.3e}, energy rel err {ee:.3e}")
    assert ew < 1e-10 and ed < BAR and ee < BAR


@pytest.mark.parametrize("form", ["int8_sliced", "bf16_sliced"])
def test_sliced_forms_on_a_grid_beyond_64_blocks(form):
    """8201 points (1640 elements of the 4096-point grid's length on [-400, 400]): 132 k blocks of
    64 points, so the active-block lists of the contraction and of the direct G tiles walk three
    words of 64 flags, and the active blocks (|x| < 55) straddle the first two.  Against the
    oracle, to the same bars as the 4096-point system."""
    w = small("long_8201", -400.0, 400.0, 1640, 6, 4, 8)
    assert w.nG == 8201
    s, x, wts = oracle_system(w)
    ctx, xg, wg = wl.make_context(w)
    state = wl.seeded_state(xg, wg, w.N, w.M, seed=5)
    _, orbs = wl.split_state(state, w.nC, w.nG, w.M)
    print(f"\n[{form}: nG={w.nG} N={w.N} M={w.M} nC={w.nC}]", flush=True)
    try:
        ctx.set_precision(hip.PREC_INT8_SLICED if form == "int8_sliced" else hip.PREC_BF16_SLICED)
        ew = rel(ctx.meanfield_w(orbs), s.meanfield_w(orbs))
        ed = state_rel(ctx.time_derivative(state), s.time_derivative(state), w.nC, wts, w.M)
        ee = abs(ctx.energy(state) - s.energy(state)) / abs(s.energy(state))
    finally:
        ctx.close()
    print(f"{form}, 8201 points: W rel err {ew:.3e}, RHS rel err {ed:.3e}, energy rel err {ee:.3e}")
    assert ew < 1e-10 and ed < BAR and ee < BAR
