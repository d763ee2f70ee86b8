"""Labelled interval timer, mirroring qubo_lib's PerformanceTimer (src/qubo_lib/utils/perf_timer.py:4-39).

`times` is a list of (elapsed_ms_since_previous_pick, label) and `get_time_from_start`
their sum (perf_timer.py:34-38).  Host wall time only: device time is measured with HIP
events inside libqbx.so and reported separately (Solution.kernel_ms).
"""
from __future__ import annotations

import time


class PerformanceTimer:
    def __init__(self, use_gpu: bool = False):
        self.times = []
        self.use_gpu = use_gpu
        self.start_time = time.perf_counter()

    def pick_time(self, label=None):
        now = time.perf_counter()
        self.times.append(((now - self.start_time) * 1000.0, label))
        self.start_time = now

    def get_time_from_start(self) -> float:
        return sum(t for t, _ in self.times)
