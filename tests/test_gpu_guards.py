"""Host-side argument checks of the contraction launchers.

The k-split partial sums of the mean-field contraction live in a buffer of kMaxWSplit = 8 planes,
and the kernels are instantiated for 1..10 (FP64) or 1..5 (int8) column tiles.  A k split outside
{1, 2, 4, 8}, or one that does not divide the kernel's k blocks, is refused on the host with
CODY_ERR_ARG before anything is launched; the context keeps its split and still matches the
oracle afterwards.
"""
import pytest

from codyfortran_amd import hip
from codyfortran_amd import workloads as wl
from tests.common import oracle_system, state_rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def system():
    w = wl.WORKLOADS["mctdhf_1d_401"]
    s, x, wts = oracle_system(w)
    ctx, xg, wg = wl.make_context(w)
    state = wl.seeded_state(xg, wg, w.N, w.M, seed=21)
    yield w, s, ctx, state, wts
    ctx.close()


@pytest.mark.parametrize("split", [16, 3, 12])
def test_bad_k_split_is_refused_before_launch(system, split):
    w, s, ctx, state, wts = system
    before = ctx.autotune()  # the timed choice, a valid split
    with pytest.raises(hip.CodyError) as err:
        ctx.autotune(split)
    assert "k split" in str(err.value)
    assert "cody error 1" in str(err.value)  # CODY_ERR_ARG
    # the context kept its split and its right-hand side
    assert ctx.autotune(before) == before
    d = ctx.time_derivative(state)
    assert state_rel(d, s.time_derivative(state), w.nC, wts, w.M) < 1e-12


def test_every_valid_split_matches_the_oracle(system):
    w, s, ctx, state, wts = system
    ref = s.time_derivative(state)
    try:
        for split in (1, 2, 4, 8):
            assert ctx.autotune(split) == split
            assert state_rel(ctx.time_derivative(state), ref, w.nC, wts, w.M) < 1e-12
    finally:
        ctx.autotune()
