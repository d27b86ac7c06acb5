"""cody_project_orbs_dev_noalloc (the projector with static scratch, no allocation) against the
allocating entry point cody_project_orbs, through the C ABI: two calls in a row, the second with a
smaller input (fewer points and orbitals, so it reads only partial sums it wrote itself), each
within 1e-13 relative of the existing entry point (the two share the overlap kernel and the
summation order of the partial sums); sizes above the static scratch are refused on the host.
"""
import numpy as np
import pytest

from codyfortran_amd import hip
from codyfortran_amd import workloads as wl
from tests.common import rel

pytestmark = pytest.mark.gpu


def _case(w, seed):
    ctx, x, wts = wl.make_context(w)
    rng = np.random.default_rng(seed)
    orbs = wl.orthonormalize(rng.standard_normal((w.nG, w.M)) + 1j * rng.standard_normal((w.nG, w.M)), wts)
    dorbs = np.asfortranarray(rng.standard_normal((w.nG, w.M)) + 1j * rng.standard_normal((w.nG, w.M)))
    return ctx, wts, np.asfortranarray(orbs), dorbs


def _dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a).view(np.float64).copy()).to(torch.device("cuda:0"))


def test_noalloc_projector_matches_allocating_entry_point():
    import torch

    stream = torch.cuda.current_stream().cuda_stream
    # first call 4096 points x 8 orbitals, then 401 points x 4 orbitals
    for w, seed in ((wl.WORKLOADS["mctdhf_1d_4096"], 1), (wl.WORKLOADS["mctdhf_1d_401"], 2)):
        print(f"\n[no-alloc projector: nG={w.nG} nO={w.M}]", flush=True)
        ctx, wts, orbs, dorbs = _case(w, seed)
        ref = ctx.project_orbs(orbs, dorbs)
        ctx.close()
        d_w = _dev(wts)
        d_o = _dev(orbs.T)  # column-major nG x nO = orbital after orbital
        d_f = _dev(dorbs.T)
        hip.project_orbs_dev_noalloc(w.nG, w.M, d_w.data_ptr(), d_o.data_ptr(), d_f.data_ptr(), stream)
        torch.cuda.synchronize()
        got = d_f.cpu().numpy().view(np.complex128).reshape(w.M, w.nG).T
        err = rel(got, ref)
        print(f"  rel err vs cody_project_orbs {err:.2e}")
        assert err < 1e-13
        # the result is orthogonal to the orbitals in the weights
        ov = orbs.conj().T @ (got * wts[:, None])
        assert np.max(np.abs(ov)) < 1e-12 * np.sqrt(np.sum(np.abs(got) ** 2 * wts[:, None]))


def test_noalloc_projector_refuses_sizes_beyond_its_scratch():
    import torch

    d = torch.zeros(16, dtype=torch.float64, device=torch.device("cuda:0"))
    p = d.data_ptr()
    with pytest.raises(hip.CodyError, match="cody error 1"):
        hip.project_orbs_dev_noalloc(65537, 4, p, p + 64, p + 128)
    with pytest.raises(hip.CodyError, match="cody error 1"):
        hip.project_orbs_dev_noalloc(100, 13, p, p + 64, p + 128)
    with pytest.raises(hip.CodyError, match="alias"):
        hip.project_orbs_dev_noalloc(4, 1, p, p + 64, p + 64)
