"""CPU checks of the C ABI: the library loads, exports every declared entry point, and its
host-side grid helpers agree with the oracle.  No compute is launched (no GPU here)."""
import ctypes as C

import numpy as np
import pytest

import oracle
from codyfortran_amd import hip


def test_library_exports_every_declared_symbol():
    lib = hip.load()
    syms = hip.header_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(lib, s), s


@pytest.mark.parametrize("nE,nLoc,lo,hi", [(819, 6, False, False), (100, 5, False, False),
                                           (10, 21, True, False), (7, 4, True, True)])
def test_fedvr_grid_matches_oracle(nE, nLoc, lo, hi):
    x1, w1 = hip.fedvr_grid(-3.0, 5.0, nE, nLoc, lo, hi)
    x2, w2 = oracle.fedvr_grid(-3.0, 5.0, nE, nLoc, lo, hi)
    np.testing.assert_allclose(x1, x2, rtol=0, atol=1e-13)
    np.testing.assert_allclose(w1, w2, rtol=1e-13)


@pytest.mark.parametrize("nLoc", [2, 3, 5, 6, 9, 21])
def test_fedvr_kref_matches_oracle(nLoc):
    K1 = hip.fedvr_kref(nLoc)
    K2, _ = oracle.fedvr_kref(nLoc)
    np.testing.assert_allclose(K1, K2, rtol=1e-11, atol=1e-11 * np.abs(K2).max())


def test_invalid_arguments_are_reported():
    lib = hip.load()
    h = C.c_void_p()
    # 13 orbitals exceed the kernels' limit: reported, not crashed (checked before any device call)
    rc = lib.cody_create(C.byref(h), 0, 100, 2, 13, 1e-8)
    assert rc == 4 and not h.value
    assert b"nOrbs" in lib.cody_last_error()
    assert lib.cody_fedvr_npoints(0, 5, 0, 0) == -1
    with pytest.raises(hip.CodyError):
        hip.fedvr_grid(1.0, -1.0, 4, 5)
