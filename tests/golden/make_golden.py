"""Writes tests/golden/mctdhf_1d_401.npz: the FP64 oracle's right-hand side, energy, dipole and
mean-field potentials for the seeded state of the 401-point workload (100 FEDVR elements x 5 GLL
points on [-100, 100], N = 2 fermions in M = 4 orbitals), plus the state itself.

The oracle is pinned against the reference's own published values (tests/test_oracle_he1d.py)
and against the Dirac-Frenkel and many-body checks (tests/test_oracle_tdvp.py,
tests/test_oracle_dipole.py); the fixture freezes its output so that a change of the oracle or of
the HIP path shows up against data that does not move.  Run from the repository root:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from codyfortran_amd import workloads as wl  # noqa: E402
from tests.common import oracle_system  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mctdhf_1d_401.npz")


def main():
    w = wl.WORKLOADS["mctdhf_1d_401"]
    s, x, wts = oracle_system(w)
    state = wl.seeded_state(x, wts, w.N, w.M, seed=1234)
    _, orbs = wl.split_state(state, w.nC, w.nG, w.M)
    np.savez_compressed(OUT, x=x, w=wts, state=state, rhs=s.time_derivative(state),
                        rhs_nonlin=s.time_derivative_nonlin(state), energy=np.array(s.energy(state)),
                        dipole=np.array(s.dipole(state)), W=s.meanfield_w(orbs))
    print("wrote", OUT)


if __name__ == "__main__":
    main()
