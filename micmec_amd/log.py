"""Minimal screen logger and section timer (stand-in for molmod's ScreenLog/TimerGroup used by
micmec/log.py:29-85).

The level comes from the environment variable MICMEC_AMD_LOG (silent, warning, low, medium,
high; default silent). Screen-log hooks are only due when the level is at least medium, so a
default run never forces a device synchronisation for printing.
"""

import os
import sys
import time
from contextlib import contextmanager

__all__ = ["log", "timer"]

_LEVELS = {"silent": 0, "warning": 1, "low": 2, "medium": 3, "high": 4, "debug": 5}


class ScreenLog(object):
    def __init__(self, level=None, stream=None):
        self.set_level(level or os.environ.get("MICMEC_AMD_LOG", "silent"))
        self.stream = stream or sys.stdout

    def set_level(self, level):
        self.level = _LEVELS[level] if isinstance(level, str) else int(level)

    do_warning = property(lambda self: self.level >= 1)
    do_low = property(lambda self: self.level >= 2)
    do_medium = property(lambda self: self.level >= 3)
    do_high = property(lambda self: self.level >= 4)

    def __call__(self, *words):
        self.stream.write(" ".join(str(w) for w in words).replace("&", " ") + "\n")

    def warn(self, *words):
        if self.do_warning:
            self("WARNING:", *words)

    def hline(self, char="~"):
        self(char * 79)


class TimerGroup(object):
    def __init__(self):
        self.totals = {}

    @contextmanager
    def section(self, label):
        t0 = time.perf_counter()
        try:
            yield
        finally:
            self.totals[label] = self.totals.get(label, 0.0) + time.perf_counter() - t0


log = ScreenLog()
timer = TimerGroup()
