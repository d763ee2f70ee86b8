"""qubo_lib_amd -- MI355X (gfx950) exhaustive QUBO solver with qubo_lib's interface.

    from qubo_lib_amd.problems import RandomQUBO, TSP, MaxCut, Just_Q, ReducedProblem
    from qubo_lib_amd.solvers import QBF, QB_F1, Parallel
    sol = QBF().solve(RandomQUBO.init_random_instance(seed=42, N=37), device="cuda")

Solvers accept the reference's problem objects as well as the mirrors in
qubo_lib_amd.problems.  On device='cuda' the search runs in hand-written HIP kernels
(libqbx.so, C ABI in include/qbx.h), never falling back to the CPU; device='cpu' (the
reference's default) runs the same lane walk built for the host (libqbx_cpu.so,
include/qbx_cpu.h), with the same answers.
"""
import logging

from .solution import Solution
from .state_vectors import StateVectors

logging.getLogger(__name__).addHandler(logging.NullHandler())

__all__ = ["Solution", "StateVectors"]
__version__ = "0.1.0"
