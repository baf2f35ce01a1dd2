"""ubolt_amd: an MI355X-native implementation of UBOLT's matrix-free
linear-Boltzmann transport operator and its preconditioner applies.

The compute path is libubolt_amd.so (hand-written HIP kernels for gfx950 behind
the C ABI of include/ubt.h); this package is the Python binding used by the
tests, the benchmark and the sensitivity harness.
"""
from .capi import UbtError, FP64, FP32  # noqa: F401
from .core import (Quadrature, Mesh, Disc, Operator, BlockInverse, DSA, Sweep, Solver, Batch, BatchSolver,  # noqa: F401
                   lqn_octant)
