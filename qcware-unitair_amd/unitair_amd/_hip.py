"""Minimal ctypes access to the HIP runtime torch already loaded (events only).

Used to hand hipEvent_t handles to ua_run_circuit for per-gate timing.
"""
from __future__ import annotations

import ctypes

import torch  # noqa: F401  (loads libamdhip64.so.7 first)

_rt = None


def runtime() -> ctypes.CDLL:
    global _rt
    if _rt is None:
        rt = ctypes.CDLL("libamdhip64.so.7")
        rt.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        rt.hipEventCreate.restype = ctypes.c_int
        rt.hipEventDestroy.argtypes = [ctypes.c_void_p]
        rt.hipEventDestroy.restype = ctypes.c_int
        rt.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p,
                                           ctypes.c_void_p]
        rt.hipEventElapsedTime.restype = ctypes.c_int
        rt.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        rt.hipEventSynchronize.restype = ctypes.c_int
        _rt = rt
    return _rt


class EventPool:
    """2*n hipEvent_t handles for per-gate (start, end) timing.  shared=True:
    gate i's start is gate i-1's end (n + 1 events, recorded once each), so a
    gate's time also holds the gap before it."""

    def __init__(self, n_gates: int, shared: bool = False):
        rt = runtime()
        self.n = n_gates
        self.handles = (ctypes.c_void_p * (2 * n_gates))()
        count = n_gates + 1 if shared else 2 * n_gates
        made = []
        for _ in range(count):
            h = ctypes.c_void_p()
            if rt.hipEventCreate(ctypes.byref(h)) != 0:
                raise RuntimeError("hipEventCreate failed")
            made.append(h.value)
        for i in range(n_gates):
            self.handles[2 * i] = made[i] if shared else made[2 * i]
            self.handles[2 * i + 1] = made[i + 1] if shared else made[2 * i + 1]
        self._made = made
        self._alive = True

    def elapsed_ms(self):
        """Per-gate durations (ms) of the last recording; synchronises."""
        rt = runtime()
        out = []
        ms = ctypes.c_float()
        rt.hipEventSynchronize(self.handles[2 * self.n - 1])
        for i in range(self.n):
            if rt.hipEventElapsedTime(ctypes.byref(ms), self.handles[2 * i],
                                      self.handles[2 * i + 1]) != 0:
                raise RuntimeError("hipEventElapsedTime failed")
            out.append(ms.value)
        return out

    def __del__(self):
        if getattr(self, "_alive", False) and _rt is not None:
            for h in self._made:
                _rt.hipEventDestroy(ctypes.c_void_p(h))
            self._alive = False
