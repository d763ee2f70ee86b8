from .parallel import Parallel
from .qbf import QBF
from .qbf1 import QB_F1

__all__ = ["QBF", "QB_F1", "Parallel"]
