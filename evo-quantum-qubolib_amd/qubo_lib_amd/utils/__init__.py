from .perf_timer import PerformanceTimer

__all__ = ["PerformanceTimer"]
