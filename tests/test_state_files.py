"""Restart files in CodyFortran's on-disk format (coeffs.in, orbBB_II.in, rdm2BtXX_YY.in).

The reference writes them with Fortran unformatted stream I/O (src/Utils/M_Utils_DataStorage.f90:
20-105): raw complex128 values, no header, names with two-digit 1-based indices
(src/Orbs/S_Orbs.f90:251-275, src/Coeffs/S_Coeffs.f90:167-214).  No reference-written file
exists to compare against, so the byte layout is pinned twice: against numpy's raw complex128
files and against a Fortran program (tests/fortran/T_StateFiles.f90) that writes and reads the
files with stream I/O itself.  The library entry points touch only host memory; no GPU is used.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
from codyfortran_amd import build as cbuild
from codyfortran_amd import hip

EXE = os.path.join(cbuild.BIN, "t_state_files")


def _state(nC, nG, n_orbs, seed=0):
    rng = np.random.default_rng(seed)
    n = nC + nG * int(np.sum(n_orbs))
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def test_fortran_stream_io_reads_and_writes_the_same_bytes(tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("run __graft_entry__.build() first")
    (tmp_path / "f").mkdir()
    (tmp_path / "g").mkdir()
    r = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "T_StateFiles PASS" in r.stdout


@pytest.mark.parametrize("n_orbs", [4, [3, 2]])
def test_files_are_raw_complex128_with_reference_names(tmp_path, n_orbs):
    nC, nG = 10, 13
    st = _state(nC, nG, n_orbs, seed=1)
    hip.save_state_files(str(tmp_path), st, nC, nG, n_orbs)
    counts = np.atleast_1d(n_orbs)
    names = ["coeffs.in"] + [f"orb{b + 1:02d}_{i + 1:02d}.in" for b, n in enumerate(counts) for i in range(n)]
    assert sorted(os.listdir(tmp_path)) == sorted(names)
    np.testing.assert_array_equal(np.fromfile(tmp_path / "coeffs.in", dtype="<c16"), st[:nC])
    o = nC
    for name in names[1:]:
        assert os.path.getsize(tmp_path / name) == 16 * nG
        np.testing.assert_array_equal(np.fromfile(tmp_path / name, dtype="<c16"), st[o:o + nG])
        o += nG
    np.testing.assert_array_equal(hip.load_state_files(str(tmp_path), nC, nG, n_orbs), st)


def test_save_replaces_and_load_accepts_longer_files(tmp_path):
    nC, nG = 3, 5
    st = _state(nC, nG, 2, seed=2)
    np.zeros(100, dtype=np.complex128).tofile(tmp_path / "coeffs.in")  # older, longer file
    hip.save_state_files(str(tmp_path), st, nC, nG, 2)
    assert os.path.getsize(tmp_path / "coeffs.in") == 16 * nC  # replaced, as status="replace"
    # a longer file is read from its start (LoadData reads the bytes it needs)
    np.concatenate([st[nC:nC + nG], np.ones(4)]).astype(np.complex128).tofile(tmp_path / "orb01_01.in")
    np.testing.assert_array_equal(hip.load_state_files(str(tmp_path), nC, nG, 2), st)


def test_load_fails_on_missing_or_short_files(tmp_path):
    nC, nG = 3, 5
    st = _state(nC, nG, 2, seed=3)
    hip.save_state_files(str(tmp_path), st, nC, nG, 2)
    with pytest.raises(hip.CodyError, match="more data requested than stored"):
        hip.load_state_files(str(tmp_path), nC, nG + 1, 2)
    os.remove(tmp_path / "orb01_02.in")
    with pytest.raises(hip.CodyError, match="orb01_02.in"):
        hip.load_state_files(str(tmp_path), nC, nG, 2)
    with pytest.raises(hip.CodyError, match="size"):
        hip.save_state_files(str(tmp_path), st[:-1], nC, nG, 2)


def test_two_rdm_blocks_per_body_type_pair(tmp_path):
    n_orbs = [3, 2]
    nO = 5
    rng = np.random.default_rng(4)
    r2 = rng.standard_normal((nO,) * 4) + 1j * rng.standard_normal((nO,) * 4)
    r2 = np.asfortranarray(r2)
    hip.save_two_rdm(str(tmp_path), r2, n_orbs)
    lo = [0, 3]
    for b1 in range(2):
        for b2 in range(2):
            s1 = slice(lo[b1], lo[b1] + n_orbs[b1])
            s2 = slice(lo[b2], lo[b2] + n_orbs[b2])
            blk = r2[s1, s2, s1, s2]
            got = np.fromfile(tmp_path / f"rdm2Bt{b1 + 1:02d}_{b2 + 1:02d}.in", dtype="<c16")
            np.testing.assert_array_equal(got, blk.ravel(order="F"))


def test_two_rdm_of_one_body_type_is_the_oracle_rho2(tmp_path):
    x, w = oracle.fedvr_grid(-5.0, 5.0, 3, 5)
    s = oracle.System(x, w, [2], [4])
    rng = np.random.default_rng(5)
    C = rng.standard_normal(s.nC) + 1j * rng.standard_normal(s.nC)
    C /= np.linalg.norm(C)
    r2 = s.rdm2(C)
    hip.save_two_rdm(str(tmp_path), r2, 4)
    got = np.fromfile(tmp_path / "rdm2Bt01_01.in", dtype="<c16")
    np.testing.assert_array_equal(got, r2.ravel(order="F"))
