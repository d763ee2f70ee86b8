"""qubo_lib_amd -- MI355X (gfx950) exhaustive QUBO solver with qubo_lib's interface.

    from qubo_lib_amd.problems import RandomQUBO, TSP, MaxCut, Just_Q, ReducedProblem
    from qubo_lib_amd.solvers import QBF, QB_F1, Parallel
    sol = QBF().solve(RandomQUBO.init_random_instance(seed=42, N=37))

Solvers accept the reference's problem objects as well as the mirrors in
qubo_lib_amd.problems.  The search itself runs in hand-written HIP kernels (libqbx.so,
C ABI in include/qbx.h); there is no CPU fallback.
"""
import logging

from .solution import Solution
from .state_vectors import StateVectors

logging.getLogger(__name__).addHandler(logging.NullHandler())

__all__ = ["Solution", "StateVectors"]
__version__ = "0.1.0"
