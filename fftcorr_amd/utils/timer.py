"""Accumulating wall-clock timer with the reference's interface
(fftcorr/utils/timer.py:4-39)."""

from timeit import default_timer


class Timer:
    def __init__(self):
        self._on = False
        self._t0 = 0.0
        self._elapsed = 0.0

    def start(self):
        if self._on:
            raise ValueError("Timer already started")
        self._t0 = default_timer()
        self._on = True

    def stop(self):
        if not self._on:
            raise ValueError("Timer not started")
        self._elapsed += default_timer() - self._t0
        self._on = False

    def clear(self):
        if self._on:
            raise ValueError("Timer currently running")
        self._elapsed = 0.0

    def __enter__(self):
        self.start()
        return self

    def __exit__(self, *args):
        self.stop()

    @property
    def elapsed(self):
        return self._elapsed
