from .problem_abc import ProblemABC
from .solver_abc import SolverABC

__all__ = ["ProblemABC", "SolverABC"]
