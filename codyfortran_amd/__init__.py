"""MI355X (gfx950) MCTDHX right-hand side and Krylov propagation behind a C ABI.

The compute path is ``libcody_hip.so`` (hand-written HIP kernels, built in-tree by
``codyfortran_amd.build``); ``codyfortran_amd.hip`` binds it for Python tests and
measurement, and ``codyfortran_amd/fortran`` binds it for Fortran hosts.
"""
__all__ = ["hip", "workloads", "build"]
