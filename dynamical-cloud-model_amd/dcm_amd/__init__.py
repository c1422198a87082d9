"""Host-side Python package of the MI355X dynamical-core hot path.

`workloads` builds the seeded synthetic domains (grid metrics, base state and
initial fields) the parity tests and the benchmark run; `hip` binds the
C-ABI library ``libdcm_hip.so`` (the HIP kernels) with ctypes.
"""
