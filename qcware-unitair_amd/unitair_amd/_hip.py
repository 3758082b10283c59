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
    """2*n hipEvent_t handles for per-gate (start, end) timing."""

    def __init__(self, n_gates: int):
        rt = runtime()
        self.n = n_gates
        self.handles = (ctypes.c_void_p * (2 * n_gates))()
        for i in range(2 * n_gates):
            h = ctypes.c_void_p()
            if rt.hipEventCreate(ctypes.byref(h)) != 0:
                raise RuntimeError("hipEventCreate failed")
            self.handles[i] = h
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
            for h in self.handles:
                _rt.hipEventDestroy(h)
            self._alive = False
