"""Hierarchical wall-clock phase timer with the convention of the reference's
common/utils/timer.py:10-52: the first round is reported separately as
warm-up ("compile time") and averages are taken over the later rounds."""
from __future__ import annotations

import time
from collections import OrderedDict


class Timers:
    def __init__(self):
        self.total = OrderedDict()
        self.rounds = 0
        self.warmup = 0.0
        self._stack = []

    def __call__(self, name):
        return _Phase(self, name)

    def end_round(self):
        self.rounds += 1
        if self.rounds == 1:
            self.warmup = sum(v for k, v in self.total.items() if "/" not in k)
            for k in self.total:
                self.total[k] = 0.0

    def averages(self):
        n = max(1, self.rounds - 1)
        return {k: v / n for k, v in self.total.items()}


class _Phase:
    def __init__(self, t, name):
        self.t, self.name = t, name

    def __enter__(self):
        self.t._stack.append(self.name)
        self.key = "/".join(self.t._stack)
        self.t0 = time.perf_counter()

    def __exit__(self, *exc):
        self.t.total[self.key] = self.t.total.get(self.key, 0.0) + time.perf_counter() - self.t0
        self.t._stack.pop()
