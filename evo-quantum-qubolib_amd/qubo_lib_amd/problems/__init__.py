from .just_Q import Just_Q
from .max_cut import MaxCut
from .random_qubo import RandomQUBO
from .reduced_problem import ReducedProblem
from .tsp import TSP

__all__ = ["TSP", "RandomQUBO", "Just_Q", "MaxCut", "ReducedProblem"]
