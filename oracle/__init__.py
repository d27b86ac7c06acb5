"""FP64 CPU oracle for the MCTDHX right-hand side (test infrastructure only).

This package is the checker.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of ``bench.py`` may import it; the product path never
does.  It wraps ``oracle/build/libcody_oracle.so`` (built from
``oracle/cody_oracle.c`` by ``oracle/Makefile``), a C restatement of the
reference CodyFortran algorithm; every C function cites the reference
file:line it follows.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

RHS_FULL, RHS_LIN, RHS_NONLIN, RHS_IMAG, RHS_NONLIN_CONSISTENT = 0, 1, 2, 3, 4


def build(native: bool = False, build_dir: str | None = None) -> str:
    """Compile the oracle library; returns the path of the shared object."""
    bdir = build_dir or os.path.join(_HERE, "build")
    args = ["make", "-s", "-C", _HERE, f"BUILD={bdir}"]
    if native:
        args.append("ARCH=-march=native")
    subprocess.run(args, check=True)
    return os.path.join(bdir, "libcody_oracle.so")


def set_compaction(on: bool):
    """Zero-source compaction of the mean-field sweep (the reference's convolution skips zero
    sources); on by default.  Results are the same either way; off times the dense sweep."""
    load().orc_set_compaction(int(bool(on)))


def load(path: str | None = None):
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    so = path or os.path.join(_HERE, "build", "libcody_oracle.so")
    if not os.path.exists(so):
        so = build()
    lib = C.CDLL(so)
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    vp = C.c_void_p
    sig = {
        "orc_gll": (C.c_int, [C.c_int, C.c_double, C.c_double, dp, dp]),
        "orc_fedvr_npoints": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int]),
        "orc_fedvr_grid": (C.c_int, [C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, dp, dp]),
        "orc_fedvr_kref": (C.c_int, [C.c_int, dp, dp]),
        "orc_fedvr_d2": (C.c_int, [C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp]),
        "orc_fedvr_d1": (C.c_int, [C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp]),
        "orc_create": (vp, [C.c_int, dp, dp, C.c_int, ip, ip, C.c_double]),
        "orc_destroy": (None, [vp]),
        "orc_ncoeffs": (C.c_int, [vp]),
        "orc_norbs": (C.c_int, [vp]),
        "orc_set_kinetic_fedvr": (C.c_int, [vp, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, dp]),
        "orc_set_kinetic_fd": (C.c_int, [vp, C.c_double, dp]),
        "orc_set_potential": (C.c_int, [vp, dp]),
        "orc_set_interaction": (C.c_int, [vp, dp]),
        "orc_set_laser": (C.c_int, [vp, C.c_int, dp, C.c_double]),
        "orc_rdm1": (C.c_int, [vp, vp, vp]),
        "orc_rdm2": (C.c_int, [vp, vp, vp]),
        "orc_apply_h1": (C.c_int, [vp, vp, vp, vp]),
        "orc_apply_h2": (C.c_int, [vp, vp, vp, vp]),
        "orc_config_codes": (C.c_int, [vp, C.c_int, vp]),
        "orc_apply_single_body": (C.c_int, [vp, vp, C.c_double, vp, vp]),
        "orc_meanfield_w": (C.c_int, [vp, vp, vp]),
        "orc_set_compaction": (None, [C.c_int]),
        "orc_active_sources": (C.c_int, [vp, vp]),
        "orc_apply_correlation": (C.c_int, [vp, vp, vp, vp, vp, vp]),
        "orc_fill_h1": (C.c_int, [vp, vp, C.c_double, vp]),
        "orc_fill_h2": (C.c_int, [vp, vp, vp]),
        "orc_reg_inverse": (C.c_int, [vp, vp, vp]),
        "orc_project_orbs": (C.c_int, [vp, vp, vp]),
        "orc_time_derivative": (C.c_int, [vp, vp, C.c_double, vp]),
        "orc_time_derivative_orbs_lin": (C.c_int, [vp, vp, C.c_double, vp]),
        "orc_time_derivative_nonlin": (C.c_int, [vp, vp, C.c_double, vp]),
        "orc_energy": (C.c_double, [vp, vp, C.c_double]),
        "orc_dipole": (C.c_double, [vp, vp]),
        "orc_rk4": (C.c_int, [vp, C.c_int, vp, C.c_double, C.c_double]),
        "orc_rk1": (C.c_int, [vp, C.c_int, vp, C.c_double, C.c_double]),
        "orc_sil": (C.c_int, [vp, C.c_int, vp, C.c_double, C.c_double, C.c_int, C.c_double, C.c_int, vp]),
        "orc_strang": (C.c_int, [vp, vp, C.c_double, C.c_double, C.c_int, C.c_double, C.c_int, vp]),
        "orc_strang_split": (C.c_int, [vp, C.c_int, vp, C.c_double, C.c_double, C.c_int, C.c_double, C.c_int, vp]),
        "orc_rhs_vjp": (C.c_int, [vp, C.c_int, vp, C.c_double, vp, vp, dp]),
        "orc_dipole_grad": (C.c_int, [vp, vp, vp]),
        "orc_cheb_degree": (C.c_int, [C.c_double, C.c_double, C.c_double]),
        "orc_strang_cheb": (C.c_int, [vp, C.c_int, vp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int]),
        "orc_dipole_gradient": (C.c_int, [vp, C.c_int, vp, C.c_double, C.c_double, C.c_int, C.c_double,
                                           C.c_double, C.c_int, dp, dp]),
        "orc_orthonormalize": (C.c_int, [vp, vp]),
        "orc_normalize": (C.c_int, [vp, vp]),
        "orc_symeig": (C.c_int, [C.c_int, dp, dp, dp]),
        "orc_num_threads": (C.c_int, []),
        "orc_set_num_threads": (None, [C.c_int]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _LIB = lib
    return lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _vp(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


def _c128(a) -> np.ndarray:
    return np.require(np.asarray(a, dtype=np.complex128), requirements=["F", "A", "W"])


def _f64(a) -> np.ndarray:
    return np.require(np.asarray(a, dtype=np.float64), requirements=["F", "A"])


# ---------------------------------------------------------------- grid helpers
def gll(n: int, a: float = -1.0, b: float = 1.0):
    lib = load()
    x = np.zeros(n)
    w = np.zeros(n)
    lib.orc_gll(n, a, b, _dp(x), _dp(w))
    return x, w


def fedvr_grid(xmin, xmax, nE, nLoc, xmin_excl=False, xmax_excl=False):
    lib = load()
    n = lib.orc_fedvr_npoints(nE, nLoc, int(xmin_excl), int(xmax_excl))
    x = np.zeros(n)
    w = np.zeros(n)
    lib.orc_fedvr_grid(xmin, xmax, nE, nLoc, int(xmin_excl), int(xmax_excl), _dp(x), _dp(w))
    return x, w


def fedvr_kref(nLoc: int):
    lib = load()
    K = np.zeros((nLoc, nLoc), order="F")
    F = np.zeros((nLoc, nLoc), order="F")
    lib.orc_fedvr_kref(nLoc, _dp(K), _dp(F))
    return K, F


def fedvr_derivative(order, f, xmin, xmax, nE, nLoc, xmin_excl=False, xmax_excl=False):
    lib = load()
    f = _c128(f)
    out = np.zeros_like(f)
    fn = lib.orc_fedvr_d2 if order == 2 else lib.orc_fedvr_d1
    fn(xmin, xmax, nE, nLoc, int(xmin_excl), int(xmax_excl), _vp(f), _vp(out))
    return out


def symeig(a):
    lib = load()
    a = _f64(a)
    n = a.shape[0]
    ev = np.zeros(n)
    V = np.zeros((n, n), order="F")
    lib.orc_symeig(n, _dp(a), _dp(ev), _dp(V))
    return ev, V


def softyukawa(d, strength, a1, a2, damp):
    """Kernel s e^{-damp d} / (sqrt(d^2 + a1^2) + a2) (S_SysInteraction_Linear_SoftYukawa.f90:67-73)."""
    d = np.abs(d)
    return strength * np.exp(-damp * d) / (np.sqrt(d * d + a1 * a1) + a2)


class System:
    """Oracle handle for one MCTDHX system (single or multiple fermionic body types)."""

    def __init__(self, x, w, n_bodies, n_orbs, reg=1e-10):
        self.lib = load()
        self.x = _f64(x).copy()
        self.w = _f64(w).copy()
        nb = np.asarray(n_bodies, dtype=np.int32)
        no = np.asarray(n_orbs, dtype=np.int32)
        self.h = self.lib.orc_create(len(self.x), _dp(self.x), _dp(self.w), len(nb),
                                     nb.ctypes.data_as(C.POINTER(C.c_int)),
                                     no.ctypes.data_as(C.POINTER(C.c_int)), float(reg))
        if not self.h:
            raise ValueError("oracle: invalid system")
        self.nG = len(self.x)
        self.nC = self.lib.orc_ncoeffs(self.h)
        self.nO = self.lib.orc_norbs(self.h)
        self.nBT = len(nb)

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self.lib.orc_destroy(h)
            self.h = None

    # setup
    def set_kinetic_fedvr(self, xmin, xmax, nE, nLoc, xmin_excl=False, xmax_excl=False, mass=None):
        m = _f64(np.ones(self.nBT) if mass is None else mass)
        r = self.lib.orc_set_kinetic_fedvr(self.h, xmin, xmax, nE, nLoc, int(xmin_excl), int(xmax_excl), _dp(m))
        if r:
            raise ValueError("oracle: FEDVR grid does not match nG")

    def set_kinetic_fd(self, dx, mass=None):
        m = _f64(np.ones(self.nBT) if mass is None else mass)
        self.lib.orc_set_kinetic_fd(self.h, dx, _dp(m))

    def set_potential(self, v):
        v = _f64(v)
        self.lib.orc_set_potential(self.h, _dp(v))

    def set_interaction_matrix(self, V):
        V = np.ascontiguousarray(V, dtype=np.float64)
        self.lib.orc_set_interaction(self.h, _dp(V))

    def set_interaction_softyukawa(self, strength=1.0, a1=1.0, a2=0.0, damp=0.0):
        d = self.x[:, None] - self.x[None, :]
        self.set_interaction_matrix(softyukawa(d, strength, a1, a2, damp))

    n_laser = 0

    def set_laser(self, c, T):
        c = _f64(c)
        self.lib.orc_set_laser(self.h, len(c), _dp(c), float(T))
        self.n_laser = len(c)

    # state layout
    @property
    def n_state(self):
        return self.nC + self.nG * self.nO

    def split(self, state):
        return state[: self.nC], state[self.nC:].reshape(self.nO, self.nG).T

    # CI
    def rdm1(self, coeffs):
        c = _c128(coeffs)
        r = np.zeros((self.nO, self.nO), dtype=np.complex128, order="F")
        self.lib.orc_rdm1(self.h, _vp(c), _vp(r))
        return r

    def rdm2(self, coeffs):
        c = _c128(coeffs)
        n = self.nO
        r = np.zeros((n, n, n, n), dtype=np.complex128, order="F")
        self.lib.orc_rdm2(self.h, _vp(c), _vp(r))
        return r

    def apply_h1(self, coeffs, h1):
        c = _c128(coeffs)
        h = _c128(h1)
        out = np.zeros(self.nC, dtype=np.complex128)
        self.lib.orc_apply_h1(self.h, _vp(c), _vp(h), _vp(out))
        return out

    def apply_h2(self, coeffs, h2):
        c = _c128(coeffs)
        h = _c128(h2)
        out = np.zeros(self.nC, dtype=np.complex128)
        self.lib.orc_apply_h2(self.h, _vp(c), _vp(h), _vp(out))
        return out

    def config_codes(self, bt=0):
        out = np.zeros(1 << 20, dtype=np.uint32)
        n = self.lib.orc_config_codes(self.h, bt, _vp(out))
        return out[:n].copy()

    # operators
    def apply_single_body(self, orbs, t=0.0):
        o = _c128(orbs)
        d = np.zeros_like(o)
        h1 = np.zeros((self.nO, self.nO), dtype=np.complex128, order="F")
        self.lib.orc_apply_single_body(self.h, _vp(o), t, _vp(d), _vp(h1))
        return d, h1

    def active_sources(self, orbs):
        """Source points the mean-field sweep visits for these orbitals (all of them with the
        compaction off)."""
        return int(self.lib.orc_active_sources(self.h, _vp(_c128(orbs))))

    def meanfield_w(self, orbs):
        o = _c128(orbs)
        P = self.nO * (self.nO + 1) // 2
        W = np.zeros((self.nG, P), dtype=np.complex128, order="F")
        self.lib.orc_meanfield_w(self.h, _vp(o), _vp(W))
        return W

    def apply_correlation(self, orbs, rdm1, rdm2):
        o = _c128(orbs)
        r1 = _c128(rdm1)
        r2 = _c128(rdm2)
        d = np.zeros_like(o)
        n = self.nO
        h2 = np.zeros((n, n, n, n), dtype=np.complex128, order="F")
        self.lib.orc_apply_correlation(self.h, _vp(o), _vp(r1), _vp(r2), _vp(d), _vp(h2))
        return d, h2

    def fill_h1(self, orbs, t=0.0):
        o = _c128(orbs)
        h1 = np.zeros((self.nO, self.nO), dtype=np.complex128, order="F")
        self.lib.orc_fill_h1(self.h, _vp(o), t, _vp(h1))
        return h1

    def fill_h2(self, orbs):
        o = _c128(orbs)
        n = self.nO
        h2 = np.zeros((n, n, n, n), dtype=np.complex128, order="F")
        self.lib.orc_fill_h2(self.h, _vp(o), _vp(h2))
        return h2

    def reg_inverse(self, rdm1):
        r = _c128(rdm1)
        out = np.zeros((self.nO, self.nO), dtype=np.complex128, order="F")
        self.lib.orc_reg_inverse(self.h, _vp(r), _vp(out))
        return out

    def project_orbs(self, orbs, dorbs):
        o = _c128(orbs)
        d = _c128(dorbs).copy(order="F")
        self.lib.orc_project_orbs(self.h, _vp(o), _vp(d))
        return d

    # method
    def _rhs(self, fn, state, t):
        s = _c128(state)
        d = np.zeros_like(s)
        fn(self.h, _vp(s), float(t), _vp(d))
        return d

    def time_derivative(self, state, t=0.0):
        return self._rhs(self.lib.orc_time_derivative, state, t)

    def time_derivative_orbs_lin(self, state, t=0.0):
        return self._rhs(self.lib.orc_time_derivative_orbs_lin, state, t)

    def time_derivative_nonlin(self, state, t=0.0):
        return self._rhs(self.lib.orc_time_derivative_nonlin, state, t)

    def energy(self, state, t=0.0):
        s = _c128(state)
        return self.lib.orc_energy(self.h, _vp(s), float(t))

    def dipole(self, state):
        s = _c128(state)
        return self.lib.orc_dipole(self.h, _vp(s))

    # integrators (in place on a copy; returned)
    def rk4(self, state, t0, t1, rhs=RHS_FULL):
        s = _c128(state).copy()
        self.lib.orc_rk4(self.h, rhs, _vp(s), t0, t1)
        return s

    def rk1(self, state, t0, t1, rhs=RHS_FULL):
        s = _c128(state).copy()
        self.lib.orc_rk1(self.h, rhs, _vp(s), t0, t1)
        return s

    def sil(self, state, t0, t1, krylov_dim=32, tol=1e-10, max_restarts=8, metric=None, rhs=RHS_LIN):
        s = _c128(state).copy()
        m = None if metric is None else _f64(metric)
        ok = self.lib.orc_sil(self.h, rhs, _vp(s), t0, t1, krylov_dim, tol, max_restarts,
                              None if m is None else _vp(m))
        return s, bool(ok)

    def strang(self, state, t0, t1, krylov_dim=32, tol=1e-10, max_restarts=8, metric=None, consistent=False):
        """Strang step; consistent=True integrates B = full - linear part instead of the
        reference's nonlinear part."""
        s = _c128(state).copy()
        m = None if metric is None else _f64(metric)
        ok = self.lib.orc_strang_split(self.h, RHS_NONLIN_CONSISTENT if consistent else RHS_NONLIN, _vp(s), t0, t1,
                                       krylov_dim, tol, max_restarts, None if m is None else _vp(m))
        return s, bool(ok)

    # adjoint (one body type); gradients in the metric of metric_weights()
    def rhs_vjp(self, state, lam, t=0.0, rhs=RHS_FULL):
        """(J^T lam, <lam, df/dE>) of the right-hand side `rhs` at (state, t)."""
        s = _c128(state)
        l = _c128(lam)
        g = np.zeros_like(s)
        eb = C.c_double(0.0)
        if self.lib.orc_rhs_vjp(self.h, rhs, _vp(s), float(t), _vp(l), _vp(g), C.byref(eb)):
            raise ValueError("oracle: the adjoint covers one body type and the FULL/LIN/NONLIN parts")
        return g, eb.value

    def dipole_grad(self, state):
        s = _c128(state)
        g = np.zeros_like(s)
        if self.lib.orc_dipole_grad(self.h, _vp(s), _vp(g)):
            raise ValueError("oracle: the adjoint covers one body type")
        return g

    def cheb_degree(self, tau, lo, hi):
        return int(self.lib.orc_cheb_degree(float(tau), float(lo), float(hi)))

    def strang_cheb(self, state, t0, t1, lo, hi, K, consistent=False):
        s = _c128(state).copy()
        if self.lib.orc_strang_cheb(self.h, RHS_NONLIN_CONSISTENT if consistent else RHS_NONLIN, _vp(s),
                                    float(t0), float(t1), float(lo), float(hi), int(K)):
            raise ValueError("oracle: invalid Chebyshev Strang step")
        return s

    def dipole_gradient(self, state, t0, dt, n_steps, lo, hi, K, consistent=False):
        """(<x>(T), dJ/dc, final state) after n_steps Chebyshev-Strang steps."""
        s = _c128(state).copy()
        J = C.c_double(0.0)
        g = np.zeros(max(self.n_laser, 1))
        if self.lib.orc_dipole_gradient(self.h, RHS_NONLIN_CONSISTENT if consistent else RHS_NONLIN, _vp(s),
                                        float(t0), float(dt), int(n_steps), float(lo), float(hi), int(K),
                                        C.byref(J), _dp(g)):
            raise ValueError("oracle: invalid gradient request")
        return J.value, g[: self.n_laser].copy(), s

    def orthonormalize(self, orbs):
        o = _c128(orbs).copy(order="F")
        if self.lib.orc_orthonormalize(self.h, _vp(o)):
            raise ValueError("oracle: linearly dependent orbitals")
        return o

    def normalize(self, coeffs):
        c = _c128(coeffs).copy()
        self.lib.orc_normalize(self.h, _vp(c))
        return c

    def metric_weights(self):
        """SIL metric: unit weights on the CI block, grid weights on each orbital block."""
        return np.concatenate([np.ones(self.nC), np.tile(self.w, self.nO)])
