"""The Fortran host path: a program written against the reference's Method/Propagator procedure
pointers fabricates the GPU leaf and must reproduce the oracle (RHS, energy, split-step states)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from codyfortran_amd import build as cbuild
from codyfortran_amd import hip
from codyfortran_amd import workloads as wl
from tests.common import oracle_system, state_rel

EXE = os.path.join(cbuild.BIN, "t_mctdhx_hip")
SPLIT_EXE = os.path.join(cbuild.BIN, "t_mctdhx_split")


def test_fortran_driver_is_built_against_the_library():
    assert os.path.exists(EXE), "run __graft_entry__.build() first"
    out = subprocess.run(["ldd", EXE], capture_output=True, text=True).stdout
    assert "libcody_hip.so" in out


@pytest.mark.gpu
def test_fortran_driver_matches_oracle():
    w = wl.WORKLOADS["mctdhf_1d_401"]
    s, x, wts = oracle_system(w)
    state = wl.seeded_state(x, wts, w.N, w.M, seed=11)
    with tempfile.TemporaryDirectory() as d:
        inp = os.path.join(d, "state.in")
        state.astype(np.complex128).tofile(inp)
        r = subprocess.run([EXE, inp, d], capture_output=True, text=True, timeout=120)
        print(r.stdout, r.stderr)
        assert r.returncode == 0
        dstate = np.fromfile(os.path.join(d, "dstate.bin"), dtype=np.complex128)
        pieces = np.fromfile(os.path.join(d, "pieces.bin"), dtype=np.complex128)
        state5 = np.fromfile(os.path.join(d, "state5.bin"), dtype=np.complex128)
    vals = dict(line.split() for line in r.stdout.splitlines() if len(line.split()) == 2)
    assert state_rel(dstate, s.time_derivative(state), w.nC, wts, w.M) < 1e-12
    # the operator pieces through the leaf's reference-interface procedures
    import oracle

    nG, M = w.nG, w.M
    C, orbs = wl.split_state(state, w.nC, nG, M)
    o = 0
    kin, o = pieces[o:o + nG], o + nG
    pot, o = pieces[o:o + nG], o + nG
    proj, o = pieces[o:o + nG * M].reshape(M, nG).T, o + nG * M
    corr, o = pieces[o:o + nG * M].reshape(M, nG).T, o + nG * M
    h2f = pieces[o:o + M ** 4].reshape((M,) * 4, order="F")
    vext = -w.pot_charge / (np.sqrt(x ** 2 + w.pot_soft1 ** 2) + w.pot_soft2)
    tk = s.apply_single_body(orbs, t=-1.0)[0][:, 0] - vext * orbs[:, 0]
    assert np.linalg.norm(kin - tk) < 1e-12 * np.linalg.norm(tk)
    V = oracle.softyukawa(x[:, None] - x[None, :], w.int_strength, w.int_soft1, w.int_soft2, 0.0)
    pr = V @ (np.conj(orbs[:, 0]) * orbs[:, 1] * wts)
    assert np.linalg.norm(pot - pr) < 1e-13 * np.linalg.norm(pr)
    pj = s.project_orbs(orbs, x[:, None] * orbs)
    assert np.linalg.norm(proj - pj) < 1e-13 * np.linalg.norm(pj)
    co, h2o = s.apply_correlation(orbs, s.rdm1(C), s.rdm2(C))
    assert np.linalg.norm(corr - co) < 1e-12 * np.linalg.norm(co)
    assert np.linalg.norm(h2f - h2o) < 1e-12 * np.linalg.norm(h2o)
    e0 = s.energy(state)
    assert abs(float(vals["energy0"]) - e0) < 1e-12 * abs(e0)
    so = state.copy()
    t = 0.0
    for _ in range(5):
        so, _ = s.strang(so, t, t + 0.01, krylov_dim=16, tol=1e-12, max_restarts=8,
                         metric=s.metric_weights())
        t += 0.01
    assert state_rel(state5, so, w.nC, wts, w.M) < 1e-10
    assert abs(float(vals["energy5"]) - s.energy(so, t)) < 1e-10 * abs(e0)
    assert abs(float(vals["dipole5"]) - s.dipole(so)) < 1e-10
    # the leaf's dipole gradient against the oracle's adjoint on the same Chebyshev-Strang map
    c = np.array([0.4, -0.3, 0.2])
    s.set_laser(c, 0.5)
    J_o, g_o, _ = s.dipole_gradient(state, 0.0, 0.05, 10, float(vals["chebLo"]), float(vals["chebHi"]),
                                    int(vals["chebK"]))
    g_f = np.array([float(vals[f"grad{k}"]) for k in (1, 2, 3)])
    assert abs(float(vals["gradJ"]) - J_o) < 1e-11 * max(1.0, abs(J_o))
    assert np.linalg.norm(g_f - g_o) < 1e-9 * np.linalg.norm(g_o)


def test_split_chain_refuses_foreign_callback():
    """The device SIL element stops, before any device call, when its TimeDerivative is not the
    GPU leaf's linear part (a device integrator cannot call a host callback)."""
    assert os.path.exists(SPLIT_EXE), "run __graft_entry__.build() first"
    r = subprocess.run([SPLIT_EXE, "--foreign-callback"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "TimeDerivative must be the GPU leaf's TimeDerivativeOrbsLin" in r.stdout + r.stderr


@pytest.mark.gpu
def test_fortran_split_chain_matches_oracle():
    """A program wired like the reference's split-step SIL driver (integrator list fabricated with
    Method_Mb_OrbBased_TimeDerivativeOrbsLin / ...CoeffsPlusOrbsNonLin, silHip + rk4Hip elements,
    the state metric on the SIL element, A(dt/2) B(dt) A(dt/2)) against the oracle: the split
    parts, ApplySingleBodyOp / ApplyInteractionOp through the method pointers, and five steps."""
    import oracle

    w = wl.WORKLOADS["mctdhf_1d_401"]
    s, x, wts = oracle_system(w)
    state = wl.seeded_state(x, wts, w.N, w.M, seed=11)  # rho1 eigenvalues 0.14, 0.86
    nG, M, nC = w.nG, w.M, w.nC
    with tempfile.TemporaryDirectory() as d:
        inp = os.path.join(d, "state.in")
        state.astype(np.complex128).tofile(inp)
        r = subprocess.run([SPLIT_EXE, inp, d], capture_output=True, text=True, timeout=120)
        print(r.stdout, r.stderr)
        assert r.returncode == 0
        parts = np.fromfile(os.path.join(d, "split_parts.bin"), dtype=np.complex128)
        chain5 = np.fromfile(os.path.join(d, "chain5.bin"), dtype=np.complex128)
    vals = dict(line.split() for line in r.stdout.splitlines() if len(line.split()) == 2)
    nS = nC + nG * M
    o = 0
    lin, o = parts[o:o + nS], o + nS
    nonlin, o = parts[o:o + nS], o + nS
    sb, o = parts[o:o + nG * M].reshape(M, nG).T, o + nG * M
    h1, o = parts[o:o + M * M].reshape(M, M).T, o + M * M
    inter, o = parts[o:o + nG * M].reshape(M, nG).T, o + nG * M
    assert o == parts.size
    assert state_rel(lin, s.time_derivative_orbs_lin(state), nC, wts, M) < 1e-12
    assert state_rel(nonlin, s.time_derivative_nonlin(state), nC, wts, M) < 1e-12
    C, orbs = wl.split_state(state, nC, nG, M)
    xo = x[:, None] * orbs
    acc = xo + s.apply_single_body(orbs)[0]
    assert np.linalg.norm(sb - acc) < 1e-13 * np.linalg.norm(acc)
    h1o = (np.conj(orbs) * wts[:, None]).T @ acc
    assert np.linalg.norm(h1 - h1o) < 1e-13 * np.linalg.norm(h1o)
    V = oracle.softyukawa(x[:, None] - x[None, :], w.int_strength, w.int_soft1, w.int_soft2, 0.0)
    p23 = V @ (np.conj(orbs[:, 1]) * orbs[:, 2] * wts)
    p41 = V @ (np.conj(orbs[:, 3]) * orbs[:, 0] * wts)
    ref = xo + (p23 + p41)[:, None] * orbs
    assert np.linalg.norm(inter - ref) < 1e-12 * np.linalg.norm(ref)
    so, t = state.copy(), 0.0
    for _ in range(5):
        so, _ = s.strang(so, t, t + 0.01, krylov_dim=16, tol=1e-12, max_restarts=8, metric=s.metric_weights())
        t += 0.01
    assert state_rel(chain5, so, nC, wts, M) < 1e-10
    e0 = s.energy(state)
    assert abs(float(vals["startEnergy"]) - e0) < 1e-12 * abs(e0)
    assert abs(float(vals["energy5"]) - s.energy(so, t)) < 1e-10 * abs(e0)
    # RK4 on the nonlinear part is not norm preserving: the drift is the oracle's
    mw = s.metric_weights()
    drift_o = abs(np.sqrt(np.sum(mw * np.abs(so) ** 2)) - np.sqrt(np.sum(mw * np.abs(state) ** 2)))
    assert abs(float(vals["normDrift"]) - drift_o) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [hip.PREC_INT8_SLICED, hip.PREC_BF16_SLICED])
def test_fortran_driver_with_sliced_contraction(precision):
    """The fabrication input's contractionPrecision selects the contraction form through the
    Fortran leaf: the right-hand side and the energy within 1e-9 of the FP64 oracle."""
    w = wl.WORKLOADS["mctdhf_1d_401"]
    s, x, wts = oracle_system(w)
    state = wl.seeded_state(x, wts, w.N, w.M, seed=11)
    with tempfile.TemporaryDirectory() as d:
        inp = os.path.join(d, "state.in")
        state.astype(np.complex128).tofile(inp)
        r = subprocess.run([EXE, inp, d, str(precision)], capture_output=True, text=True, timeout=120)
        print(r.stdout, r.stderr)
        assert r.returncode == 0
        dstate = np.fromfile(os.path.join(d, "dstate.bin"), dtype=np.complex128)
    vals = dict(line.split() for line in r.stdout.splitlines() if len(line.split()) == 2)
    err = state_rel(dstate, s.time_derivative(state), w.nC, wts, w.M)
    print(f"precision {precision}: RHS rel err {err:.3e}")
    assert err < 1e-9
    e0 = s.energy(state)
    assert abs(float(vals["energy0"]) - e0) < 1e-9 * abs(e0)
