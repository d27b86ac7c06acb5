"""Workload definitions and seeded synthetic inputs for the 1D FEDVR MCTDHF systems.

Each workload: FEDVR grid on [xmin, xmax] with nE equal elements of nLoc Gauss-Lobatto
points (both ends kept), V_ext = -2/sqrt(x^2+1), W = 1/sqrt((x-x')^2+1), N spin-polarised
fermions in M orbitals.  Seeded state: orbitals are i.i.d. complex Gaussian values times
exp(-x^2/8), Gram-Schmidt orthonormalised (twice) in the FEDVR weights; CI coefficients are
complex Gaussian, normalised.
"""
from __future__ import annotations

from dataclasses import dataclass
from math import comb

import numpy as np


@dataclass(frozen=True)
class Workload:
    name: str
    xmin: float
    xmax: float
    nE: int
    nLoc: int
    N: int
    M: int
    reg: float = 1e-8
    pot_charge: float = 2.0
    pot_soft1: float = 1.0
    pot_soft2: float = 0.0
    int_strength: float = 1.0
    int_soft1: float = 1.0
    int_soft2: float = 0.0
    dt: float = 0.02
    krylov_dim: int = 16
    steps: int = 1000

    @property
    def nG(self) -> int:
        return self.nE * (self.nLoc - 1) + 1

    @property
    def nC(self) -> int:
        return comb(self.M, self.N)

    @property
    def n_state(self) -> int:
        return self.nC + self.nG * self.M

    @property
    def pairs(self) -> int:
        return self.M * (self.M + 1) // 2


WORKLOADS = {
    # 100 elements x 5 GLL points on [-100, 100]: 401 points, N=2, M=4; SIL dim 12, dt 0.05, 200 steps
    "mctdhf_1d_401": Workload("mctdhf_1d_401", -100.0, 100.0, 100, 5, 2, 4, dt=0.05, krylov_dim=12, steps=200),
    # 819 elements x 6 GLL points on [-200, 200]: 4096 points, N=4, M=8; SIL dim 16, dt 0.02, 1000 steps
    "mctdhf_1d_4096": Workload("mctdhf_1d_4096", -200.0, 200.0, 819, 6, 4, 8, dt=0.02, krylov_dim=16, steps=1000),
}


def batched_workloads(n: int = 64, seed: int = 3, sizes=(512, 1024, 2048, 4096), orbitals=(4, 6, 8, 12)):
    """Seeded independent instances for batched evaluation: N_g drawn uniformly from `sizes`
    (realised as the nearest FEDVR size with 6 GLL points per element and both ends kept:
    511, 1026, 2046, 4096), M uniformly from `orbitals`, N = M/2 spin-polarised fermions, and one
    soft-Coulomb softening a ~ U(0.5, 2.0) for both the external potential and the interaction.
    The elements have the length of the 4096-point grid (400/819), centred on x = 0."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        ng = int(rng.choice(sizes))
        M = int(rng.choice(orbitals))
        a = float(rng.uniform(0.5, 2.0))
        nE = int(round((ng - 1) / 5))
        half = 0.5 * nE * 400.0 / 819.0
        out.append(Workload(f"batch{i}_g{nE * 5 + 1}_m{M}", -half, half, nE, 6, M // 2, M,
                            pot_soft1=a, int_soft1=a))
    return out


def orthonormalize(orbs: np.ndarray, w: np.ndarray) -> np.ndarray:
    """Modified Gram-Schmidt, two sweeps, in the metric <f|g> = sum conj(f) g w."""
    o = np.array(orbs, dtype=np.complex128, order="F", copy=True)
    for i in range(o.shape[1]):
        for _ in range(2):
            for j in range(i):
                o[:, i] -= o[:, j] * np.sum(np.conj(o[:, j]) * o[:, i] * w)
        o[:, i] /= np.sqrt(np.sum(np.abs(o[:, i]) ** 2 * w))
    return o


def seeded_state(x: np.ndarray, w: np.ndarray, N: int, M: int, seed: int = 1234):
    rng = np.random.default_rng(seed)
    nG = len(x)
    env = np.exp(-x ** 2 / 8.0)[:, None]
    orbs = (rng.standard_normal((nG, M)) + 1j * rng.standard_normal((nG, M))) * env
    orbs = orthonormalize(orbs, w)
    nC = comb(M, N)
    C = rng.standard_normal(nC) + 1j * rng.standard_normal(nC)
    C = C / np.linalg.norm(C)
    return np.concatenate([C, orbs.T.ravel()]).astype(np.complex128)


def smooth_state(x: np.ndarray, w: np.ndarray, N: int, M: int, seed: int = 1234):
    """A low-energy seeded state on the same grid: orbital m starts as x^m exp(-x^2/8) mixed by
    complex Gaussian coefficients (seeded), then Gram-Schmidt orthonormalised in the weights;
    CI coefficients as in `seeded_state`.  Unlike i.i.d. point values, these orbitals have no
    grid-scale components, so explicit integrators stay stable at ordinary step sizes."""
    rng = np.random.default_rng(seed)
    env = np.exp(-x ** 2 / 8.0)
    basis = np.stack([(x / 2.0) ** m * env for m in range(M)], axis=1)
    mix = rng.standard_normal((M, M)) + 1j * rng.standard_normal((M, M))
    orbs = orthonormalize(basis @ mix, w)
    C = rng.standard_normal(comb(M, N)) + 1j * rng.standard_normal(comb(M, N))
    C = C / np.linalg.norm(C)
    return np.concatenate([C, orbs.T.ravel()]).astype(np.complex128)


def split_state(state: np.ndarray, nC: int, nG: int, M: int):
    return state[:nC], state[nC:].reshape(M, nG).T


def make_context(wl: Workload, device: int = 0):
    """A device context set up for workload `wl` (grid, kinetic, potentials)."""
    from . import hip

    x, w = hip.fedvr_grid(wl.xmin, wl.xmax, wl.nE, wl.nLoc)
    ctx = hip.Context(len(x), wl.N, wl.M, reg=wl.reg, device=device)
    ctx.set_grid(x, w)
    ctx.set_kinetic_fedvr(wl.xmin, wl.xmax, wl.nE, wl.nLoc)
    ctx.set_potential_softyukawa([0.0], [wl.pot_charge], [wl.pot_soft1], [wl.pot_soft2], [0.0])
    ctx.set_interaction_softyukawa(wl.int_strength, wl.int_soft1, wl.int_soft2, 0.0)
    return ctx, x, w
