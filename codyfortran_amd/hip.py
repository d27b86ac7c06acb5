"""ctypes binding of ``libcody_hip.so`` (the C ABI declared in ``include/cody_mctdhx.h``).

The Python side is plumbing for tests and measurement; the product is the HIP library,
which a Fortran host binds through iso_c_binding (see ``codyfortran_amd/fortran``).
Loading fails loudly when the library has not been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcody_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "cody_mctdhx.h")

PREC_FP64, PREC_SPLIT_FP32, PREC_BF16X3, PREC_INT8_SLICED, PREC_BF16_SLICED = 0, 1, 2, 3, 4
SPLIT_REFERENCE, SPLIT_CONSISTENT = 0, 1
PART_FULL, PART_LIN, PART_NONLIN = 0, 1, 2

_lib = None


class CodyError(RuntimeError):
    pass


def header_symbols() -> list[str]:
    """Every function declared in include/cody_mctdhx.h."""
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*|void)\s+(cody_\w+)\s*\(", txt, re.M)))


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise CodyError(f"{LIB_PATH} is missing: build it with `python -m codyfortran_amd.build`")
    lib = C.CDLL(LIB_PATH)
    dp, ip, vp, i = C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_void_p, C.c_int
    d = C.c_double
    sig = {
        "cody_last_error": (C.c_char_p, []),
        "cody_fedvr_npoints": (i, [i, i, i, i]),
        "cody_fedvr_grid": (i, [d, d, i, i, i, i, dp, dp]),
        "cody_fedvr_kref": (i, [i, dp]),
        "cody_create": (i, [C.POINTER(vp), i, i, i, i, d]),
        "cody_destroy": (i, [vp]),
        "cody_sizes": (i, [vp, ip, ip]),
        "cody_get_stream": (i, [vp, C.POINTER(vp)]),
        "cody_set_grid": (i, [vp, dp, dp]),
        "cody_set_kinetic_fedvr": (i, [vp, d, d, i, i, i, i, d]),
        "cody_set_potential": (i, [vp, dp]),
        "cody_set_potential_softyukawa": (i, [vp, i, dp, dp, dp, dp, dp]),
        "cody_set_interaction_softyukawa": (i, [vp, d, d, d, d]),
        "cody_set_laser": (i, [vp, i, dp, d]),
        "cody_set_precision": (i, [vp, i]),
        "cody_set_graphs": (i, [vp, i]),
        "cody_set_split": (i, [vp, i]),
        "cody_set_source_skip": (i, [vp, i]),
        "cody_autotune": (i, [vp, C.POINTER(C.c_int)]),
        "cody_batch_create": (i, [C.POINTER(vp), i, C.POINTER(vp)]),
        "cody_batch_destroy": (i, [vp]),
        "cody_batch_time_derivative_dev": (i, [vp, C.POINTER(vp), C.POINTER(vp), d, vp]),
        "cody_batch_autotune": (i, [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(d), i, C.POINTER(C.c_int)]),
        "cody_batch_set_schedule": (i, [vp, i, d]),
        "cody_upload_state": (i, [vp, vp]),
        "cody_download_state": (i, [vp, vp]),
        "cody_state_device_ptr": (i, [vp, C.POINTER(vp)]),
        "cody_time_derivative_dev": (i, [vp, vp, vp, d, vp]),
        "cody_time_derivative": (i, [vp, vp, vp, d]),
        "cody_time_derivative_orbs_lin_dev": (i, [vp, vp, vp, d, vp]),
        "cody_time_derivative_nonlin_dev": (i, [vp, vp, vp, d, vp]),
        "cody_time_derivative_orbs_lin": (i, [vp, vp, vp, d]),
        "cody_time_derivative_nonlin": (i, [vp, vp, vp, d]),
        "cody_energy": (i, [vp, vp, d, dp]),
        "cody_dipole": (i, [vp, vp, dp]),
        "cody_dipole_dev": (i, [vp, vp, vp, vp]),
        "cody_meanfield_w": (i, [vp, vp, vp]),
        "cody_apply_single_body": (i, [vp, vp, d, vp, vp]),
        "cody_apply_single_body_acc": (i, [vp, vp, d, vp, vp]),
        "cody_apply_interaction": (i, [vp, vp, i, i, d, vp]),
        "cody_densities": (i, [vp, vp, vp, vp, vp]),
        "cody_sil": (i, [vp, d, d, i, d, i, i, ip]),
        "cody_rk4": (i, [vp, d, d, i]),
        "cody_strang": (i, [vp, d, d, i, d, i, i, ip]),
        "cody_profile_rhs": (i, [vp, i, i, dp, C.c_char_p, i, ip]),
        "cody_profile_contraction": (i, [vp, i, dp]),
        "cody_save_state_files": (i, [C.c_char_p, i, i, i, ip, vp]),
        "cody_load_state_files": (i, [C.c_char_p, i, i, i, ip, vp]),
        "cody_save_two_rdm": (i, [C.c_char_p, i, ip, vp]),
        "cody_interaction_potential": (i, [vp, i, vp, vp]),
        "cody_apply_correlation": (i, [vp, vp, vp, vp, vp, vp]),
        "cody_project_orbs": (i, [vp, vp, vp]),
        "cody_apply_kinetic": (i, [vp, vp, vp]),
        "cody_time_derivative_vjp_dev": (i, [vp, i, vp, vp, vp, d, vp, vp]),
        "cody_dipole_grad_dev": (i, [vp, vp, vp, vp]),
        "cody_set_chebyshev": (i, [vp, d, d, dp, dp, ip]),
        "cody_strang_cheb": (i, [vp, d, d]),
        "cody_dipole_gradient": (i, [vp, d, d, i, i, dp, dp]),
        "cody_project_orbs_dev_noalloc": (i, [i, i, vp, vp, vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc: int):
    if rc != 0:
        msg = load().cody_last_error().decode(errors="replace")
        raise CodyError(f"cody error {rc}: {msg}")


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _vp(a):
    return C.c_void_p(a.ctypes.data)


def _c128(a):
    return np.require(np.asarray(a, dtype=np.complex128), requirements=["F", "A"])


# ----------------------------------------------------------------------------- grid helpers
def _orb_counts(n_orbs):
    counts = np.ascontiguousarray(np.atleast_1d(np.asarray(n_orbs, dtype=np.int32)))
    return counts, counts.ctypes.data_as(C.POINTER(C.c_int))


def save_state_files(directory, state, nC, nG, n_orbs):
    """Writes coeffs.in and orbBB_II.in (raw complex128) of the packed state [C ; orbs] into
    `directory`; n_orbs is the orbital count of each body type (an int for one body type)."""
    counts, cp = _orb_counts(n_orbs)
    st = np.ascontiguousarray(state, dtype=np.complex128)
    if st.size != nC + nG * int(counts.sum()):
        raise CodyError("save_state_files: state size does not match nC + nG * sum(n_orbs)")
    _check(load().cody_save_state_files(os.fsencode(directory), nC, nG, len(counts), cp, _vp(st)))


def load_state_files(directory, nC, nG, n_orbs):
    """Reads coeffs.in and orbBB_II.in from `directory` into a packed state [C ; orbs]."""
    counts, cp = _orb_counts(n_orbs)
    st = np.zeros(nC + nG * int(counts.sum()), dtype=np.complex128)
    _check(load().cody_load_state_files(os.fsencode(directory), nC, nG, len(counts), cp, _vp(st)))
    return st


def save_two_rdm(directory, rdm2, n_orbs):
    """Writes rdm2BtXX_YY.in for every pair of body types from the full rho2 (nO^4, column-major)."""
    counts, cp = _orb_counts(n_orbs)
    r = np.require(np.asarray(rdm2, dtype=np.complex128), requirements=["F", "A"])
    if r.size != int(counts.sum()) ** 4:
        raise CodyError("save_two_rdm: rdm2 must hold nO^4 values")
    _check(load().cody_save_two_rdm(os.fsencode(directory), len(counts), cp, _vp(r)))


def fedvr_grid(xmin, xmax, nE, nLoc, xmin_excl=False, xmax_excl=False):
    lib = load()
    n = lib.cody_fedvr_npoints(nE, nLoc, int(xmin_excl), int(xmax_excl))
    if n < 1:
        raise CodyError("bad FEDVR layout")
    x = np.zeros(n)
    w = np.zeros(n)
    _check(lib.cody_fedvr_grid(xmin, xmax, nE, nLoc, int(xmin_excl), int(xmax_excl), _dp(x), _dp(w)))
    return x, w


def fedvr_kref(nLoc):
    K = np.zeros((nLoc, nLoc), order="F")
    _check(load().cody_fedvr_kref(nLoc, _dp(K)))
    return K


def project_orbs_dev_noalloc(nG: int, nO: int, d_w: int, d_orbs: int, d_dorbs: int, stream: int = 0):
    """In-place d_dorbs <- d_dorbs - sum_i phi_i <phi_i|d_dorbs>_w on device pointers, with no
    allocation (static scratch of the library; nG <= 65536, nO <= 12)."""
    _check(load().cody_project_orbs_dev_noalloc(int(nG), int(nO), C.c_void_p(d_w), C.c_void_p(d_orbs),
                                                C.c_void_p(d_dorbs), C.c_void_p(stream or None)))


class Context:
    """One MCTDHX system on one GPU (one spin-polarised fermionic body type)."""

    def __init__(self, nG, n_bodies, n_orbs, reg=1e-10, device=0):
        self.lib = load()
        h = C.c_void_p()
        _check(self.lib.cody_create(C.byref(h), device, nG, n_bodies, n_orbs, float(reg)))
        self.h = h
        nC, nS = C.c_int(), C.c_int()
        _check(self.lib.cody_sizes(h, C.byref(nC), C.byref(nS)))
        self.nG, self.nO, self.N = nG, n_orbs, n_bodies
        self.nC, self.n_state = nC.value, nS.value

    def close(self):
        if getattr(self, "h", None):
            self.lib.cody_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self) -> int:
        s = C.c_void_p()
        _check(self.lib.cody_get_stream(self.h, C.byref(s)))
        return s.value or 0

    # ---------------------------------------------------------------- setup
    def set_grid(self, x, w):
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        self.w = np.ascontiguousarray(w, dtype=np.float64)
        _check(self.lib.cody_set_grid(self.h, _dp(self.x), _dp(self.w)))

    def set_kinetic_fedvr(self, xmin, xmax, nE, nLoc, xmin_excl=False, xmax_excl=False, mass=1.0):
        _check(self.lib.cody_set_kinetic_fedvr(self.h, xmin, xmax, nE, nLoc, int(xmin_excl),
                                               int(xmax_excl), mass))

    def set_potential(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        _check(self.lib.cody_set_potential(self.h, _dp(v)))

    def set_potential_softyukawa(self, position, charge, soft1, soft2, damp):
        arrs = [np.ascontiguousarray(np.atleast_1d(a), dtype=np.float64)
                for a in (position, charge, soft1, soft2, damp)]
        _check(self.lib.cody_set_potential_softyukawa(self.h, len(arrs[0]), *[_dp(a) for a in arrs]))

    def set_interaction_softyukawa(self, strength=1.0, soft1=1.0, soft2=0.0, damp=0.0):
        _check(self.lib.cody_set_interaction_softyukawa(self.h, strength, soft1, soft2, damp))

    n_laser = 0

    def set_laser(self, c, T):
        c = np.ascontiguousarray(c, dtype=np.float64)
        _check(self.lib.cody_set_laser(self.h, len(c), _dp(c), float(T)))
        self.n_laser = len(c)

    def set_precision(self, mode):
        _check(self.lib.cody_set_precision(self.h, int(mode)))

    def autotune(self, split: int = 0) -> int:
        """Time the FP64 contraction for each k split and keep the fastest (or set `split`);
        returns the split in use."""
        k = C.c_int(int(split))
        _check(self.lib.cody_autotune(self.h, C.byref(k)))
        return k.value

    def set_split(self, consistent: bool):
        """Nonlinear split part: the reference's (False, the default) or full minus linear (True)."""
        _check(self.lib.cody_set_split(self.h, SPLIT_CONSISTENT if consistent else SPLIT_REFERENCE))

    def set_source_skip(self, enable: bool):
        """Skip (True, the default) or contract the k blocks whose pair sources are all zero."""
        _check(self.lib.cody_set_source_skip(self.h, int(bool(enable))))

    def set_graphs(self, enable: bool):
        """True: replay captured graphs for single right-hand sides and propagator steps; False:
        launch every kernel.  Default: propagator steps replayed, right-hand sides launched."""
        _check(self.lib.cody_set_graphs(self.h, int(bool(enable))))

    # ---------------------------------------------------------------- state
    def upload_state(self, state):
        s = _c128(state)
        assert s.size == self.n_state
        _check(self.lib.cody_upload_state(self.h, _vp(s)))

    def download_state(self):
        s = np.zeros(self.n_state, dtype=np.complex128)
        _check(self.lib.cody_download_state(self.h, _vp(s)))
        return s

    def save_state(self, directory):
        """Writes the device-resident state as coeffs.in and orb01_II.in in `directory`."""
        save_state_files(directory, self.download_state(), self.nC, self.nG, self.nO)

    def load_state(self, directory):
        """Reads coeffs.in and orb01_II.in from `directory` into the device-resident state."""
        st = load_state_files(directory, self.nC, self.nG, self.nO)
        self.upload_state(st)
        return st

    def state_device_ptr(self) -> int:
        p = C.c_void_p()
        _check(self.lib.cody_state_device_ptr(self.h, C.byref(p)))
        return p.value

    # ---------------------------------------------------------------- method
    def time_derivative(self, state, t=0.0, mode="full"):
        """dState of a host state: the full right-hand side, or the split parts "lin"
        (TimeDerivativeOrbsLin) and "nonlin" (TimeDerivativeCoeffsPlusOrbsNonLin)."""
        fn = {"full": self.lib.cody_time_derivative,
              "lin": self.lib.cody_time_derivative_orbs_lin,
              "nonlin": self.lib.cody_time_derivative_nonlin}[mode]
        s = _c128(state)
        d = np.zeros(self.n_state, dtype=np.complex128)
        _check(fn(self.h, _vp(s), _vp(d), float(t)))
        return d

    def time_derivative_dev(self, d_state: int, d_dstate: int, t=0.0, stream: int = 0, mode="full"):
        fn = {"full": self.lib.cody_time_derivative_dev,
              "lin": self.lib.cody_time_derivative_orbs_lin_dev,
              "nonlin": self.lib.cody_time_derivative_nonlin_dev}[mode]
        _check(fn(self.h, C.c_void_p(d_state), C.c_void_p(d_dstate), float(t), C.c_void_p(stream or None)))

    def energy(self, state, t=0.0):
        s = _c128(state)
        e = C.c_double()
        _check(self.lib.cody_energy(self.h, _vp(s), float(t), C.byref(e)))
        return e.value

    def dipole(self, state):
        """<x> of a host state."""
        s = _c128(state)
        d = C.c_double()
        _check(self.lib.cody_dipole(self.h, _vp(s), C.byref(d)))
        return d.value

    def dipole_dev(self, d_state: int, d_out: int, stream: int = 0):
        """<x> of a device state into the double at d_out, enqueued on `stream`."""
        _check(self.lib.cody_dipole_dev(self.h, C.c_void_p(d_state), C.c_void_p(d_out), C.c_void_p(stream)))

    # ---------------------------------------------------------------- pieces
    def meanfield_w(self, orbs):
        o = _c128(orbs)
        P = self.nO * (self.nO + 1) // 2
        W = np.zeros((self.nG, P), dtype=np.complex128, order="F")
        _check(self.lib.cody_meanfield_w(self.h, _vp(o), _vp(W)))
        return W

    def apply_single_body(self, orbs, t=0.0):
        o = _c128(orbs)
        d = np.zeros((self.nG, self.nO), dtype=np.complex128, order="F")
        h1 = np.zeros((self.nO, self.nO), dtype=np.complex128, order="F")
        _check(self.lib.cody_apply_single_body(self.h, _vp(o), float(t), _vp(d), _vp(h1)))
        return d, h1

    def apply_single_body_acc(self, orbs, dorbs, t=0.0):
        """(dorbs + (T + V) orbs, h1 of the accumulated result): ApplySingleBodyOp's semantics."""
        o = _c128(orbs)
        d = _c128(dorbs).copy(order="F")
        h1 = np.zeros((self.nO, self.nO), dtype=np.complex128, order="F")
        _check(self.lib.cody_apply_single_body_acc(self.h, _vp(o), float(t), _vp(d), _vp(h1)))
        return d, h1

    def apply_interaction(self, orbs, i2, j2, dorbs, t=0.0):
        """dorbs + W_(i2 j2) orbs (0-based pair indices): ApplyInteractionOp's semantics."""
        o = _c128(orbs)
        d = _c128(dorbs).copy(order="F")
        _check(self.lib.cody_apply_interaction(self.h, _vp(o), int(i2), int(j2), float(t), _vp(d)))
        return d

    def interaction_potential(self, src):
        """pot(:, s) = sum_j K(x_i - x_j) src(x_j, s) for sources holding the weights."""
        sr = np.asfortranarray(np.asarray(src, dtype=np.complex128).reshape(self.nG, -1))
        pot = np.zeros_like(sr, order="F")
        _check(self.lib.cody_interaction_potential(self.h, sr.shape[1], _vp(sr), _vp(pot)))
        return pot

    def apply_correlation(self, orbs, rdm1, rdm2, dorbs=None):
        """(dorbs + correlation term, h2), the accumulate semantics of ApplyCorrelationOp."""
        o = _c128(orbs)
        r1 = _c128(rdm1)
        r2 = _c128(rdm2)
        d = np.zeros((self.nG, self.nO), dtype=np.complex128, order="F") if dorbs is None else _c128(dorbs).copy(order="F")
        n = self.nO
        h2 = np.zeros((n, n, n, n), dtype=np.complex128, order="F")
        _check(self.lib.cody_apply_correlation(self.h, _vp(o), _vp(r1), _vp(r2), _vp(d), _vp(h2)))
        return d, h2

    def project_orbs(self, orbs, dorbs):
        o = _c128(orbs)
        d = _c128(dorbs).copy(order="F")
        _check(self.lib.cody_project_orbs(self.h, _vp(o), _vp(d)))
        return d

    def apply_kinetic(self, orbs):
        o = _c128(orbs)
        d = np.zeros((self.nG, self.nO), dtype=np.complex128, order="F")
        _check(self.lib.cody_apply_kinetic(self.h, _vp(o), _vp(d)))
        return d

    def densities(self, coeffs):
        c = _c128(coeffs)
        n = self.nO
        r1 = np.zeros((n, n), dtype=np.complex128, order="F")
        r2 = np.zeros((n, n, n, n), dtype=np.complex128, order="F")
        ri = np.zeros((n, n), dtype=np.complex128, order="F")
        _check(self.lib.cody_densities(self.h, _vp(c), _vp(r1), _vp(r2), _vp(ri)))
        return r1, r2, ri

    # ---------------------------------------------------------------- propagation
    def sil(self, t0, t1, krylov_dim=32, tol=1e-10, max_restarts=8, metric=True):
        ok = C.c_int()
        _check(self.lib.cody_sil(self.h, t0, t1, krylov_dim, tol, max_restarts, int(metric), C.byref(ok)))
        return bool(ok.value)

    def rk4(self, t0, t1, full=False):
        _check(self.lib.cody_rk4(self.h, t0, t1, int(full)))

    def strang(self, t0, t1, krylov_dim=32, tol=1e-10, max_restarts=8, metric=True):
        ok = C.c_int()
        _check(self.lib.cody_strang(self.h, t0, t1, krylov_dim, tol, max_restarts, int(metric), C.byref(ok)))
        return bool(ok.value)

    # ---------------------------------------------------------------- adjoint
    def vjp_dev(self, part: int, d_state: int, d_lam: int, d_grad: int, t: float, d_ebar: int, stream: int = 0):
        """d_grad = J^T d_lam of the right-hand side `part` (PART_FULL / PART_LIN / PART_NONLIN) and
        the double at d_ebar = <d_lam, df/dE>, in the metric of the SIL (unit on C, w on orbitals)."""
        _check(self.lib.cody_time_derivative_vjp_dev(self.h, int(part), C.c_void_p(d_state), C.c_void_p(d_lam),
                                                     C.c_void_p(d_grad), float(t), C.c_void_p(d_ebar),
                                                     C.c_void_p(stream)))

    def dipole_grad_dev(self, d_state: int, d_grad: int, stream: int = 0):
        _check(self.lib.cody_dipole_grad_dev(self.h, C.c_void_p(d_state), C.c_void_p(d_grad), C.c_void_p(stream)))

    def set_chebyshev(self, dt: float, emax: float):
        """Fixes the Chebyshev half-steps for steps of dt and fields |E| <= emax: (lo, hi, degree)."""
        lo, hi, k = C.c_double(), C.c_double(), C.c_int()
        _check(self.lib.cody_set_chebyshev(self.h, float(dt), float(emax), C.byref(lo), C.byref(hi), C.byref(k)))
        return lo.value, hi.value, k.value

    def strang_cheb(self, t0: float, t1: float):
        _check(self.lib.cody_strang_cheb(self.h, float(t0), float(t1)))

    def dipole_gradient(self, t0: float, dt: float, n_steps: int, checkpoint_every: int = 50):
        """(<x>(t0 + n dt), dJ/dc) from the context's state, which is left at the final time."""
        J = C.c_double()
        g = np.zeros(max(self.n_laser, 1))
        _check(self.lib.cody_dipole_gradient(self.h, float(t0), float(dt), int(n_steps), int(checkpoint_every),
                                             C.byref(J), _dp(g)))
        return J.value, g[: self.n_laser].copy()

    # ---------------------------------------------------------------- measurement
    def profile_contraction(self, iters=50):
        """Average launch duration (ms) of the contraction kernel, back-to-back launches."""
        ms = C.c_double()
        _check(self.lib.cody_profile_contraction(self.h, iters, C.byref(ms)))
        return ms.value

    def profile_rhs(self, iters=20):
        n = 32
        ms = (C.c_double * n)()
        names = C.create_string_buffer(1024)
        nk = C.c_int()
        _check(self.lib.cody_profile_rhs(self.h, iters, n, ms, names, 1024, C.byref(nk)))
        labels = names.value.decode().split(";")
        return {labels[k] if k < len(labels) else f"k{k}": ms[k] for k in range(nk.value)}


class Batch:
    """Independent contexts whose right-hand sides are replayed as one captured graph."""

    def __init__(self, contexts):
        self.contexts = list(contexts)
        self.lib = load()
        arr = (C.c_void_p * len(self.contexts))(*[c.h for c in self.contexts])
        h = C.c_void_p()
        _check(self.lib.cody_batch_create(C.byref(h), len(self.contexts), arr))
        self.h = h

    def time_derivative_dev(self, d_states, d_dstates, t=0.0, stream: int = 0):
        n = len(self.contexts)
        assert len(d_states) == n and len(d_dstates) == n
        s = (C.c_void_p * n)(*d_states)
        d = (C.c_void_p * n)(*d_dstates)
        _check(self.lib.cody_batch_time_derivative_dev(self.h, s, d, float(t), C.c_void_p(stream or None)))

    # the candidate schedules of cody_batch_autotune: (parallel chains, split threshold)
    SCHEDULES = [(4, 1.0), (4, 0.7), (4, 0.5), (2, 1.0), (2, 0.5), (1, 1.0)]

    def autotune(self, d_states, d_dstates):
        """Times the candidate schedules for these state pointers and keeps the fastest; returns
        (index of the chosen schedule in SCHEDULES, ms per batched right-hand side of each
        candidate).  d_dstates are overwritten with a right-hand side at t = 0."""
        n = len(self.contexts)
        assert len(d_states) == n and len(d_dstates) == n
        s = (C.c_void_p * n)(*d_states)
        d = (C.c_void_p * n)(*d_dstates)
        k = len(self.SCHEDULES)
        ms = (C.c_double * k)()
        pick = C.c_int()
        _check(self.lib.cody_batch_autotune(self.h, s, d, ms, k, C.byref(pick)))
        return pick.value, list(ms)

    def set_schedule(self, chains: int, split_threshold: float):
        _check(self.lib.cody_batch_set_schedule(self.h, int(chains), float(split_threshold)))

    def close(self):
        if self.h:
            self.lib.cody_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
