"""Shared builders: the same seeded system in the oracle and on the GPU."""
from __future__ import annotations

import numpy as np

import oracle
from codyfortran_amd import workloads as wl


def oracle_system(w: wl.Workload):
    x, wt = oracle.fedvr_grid(w.xmin, w.xmax, w.nE, w.nLoc)
    s = oracle.System(x, wt, [w.N], [w.M], reg=w.reg)
    s.set_kinetic_fedvr(w.xmin, w.xmax, w.nE, w.nLoc)
    s.set_potential(-w.pot_charge / (np.sqrt(x ** 2 + w.pot_soft1 ** 2) + w.pot_soft2))
    s.set_interaction_softyukawa(w.int_strength, w.int_soft1, w.int_soft2, 0.0)
    return s, x, wt


def small(name, xmin, xmax, nE, nLoc, N, M, reg=1e-8):
    return wl.Workload(name, xmin, xmax, nE, nLoc, N, M, reg=reg)


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def state_rel(s_gpu, s_orc, nC, wts, M):
    """Relative error: Euclidean over C plus grid-weighted L2 over the orbitals."""
    dC = s_gpu[:nC] - s_orc[:nC]
    do = (s_gpu[nC:] - s_orc[nC:]).reshape(M, -1)
    oo = s_orc[nC:].reshape(M, -1)
    num = np.sum(np.abs(dC) ** 2) + np.sum(np.abs(do) ** 2 * wts[None, :])
    den = np.sum(np.abs(s_orc[:nC]) ** 2) + np.sum(np.abs(oo) ** 2 * wts[None, :])
    return float(np.sqrt(num / den))
