"""Minimal ctypes access to the HIP runtime torch already loaded.

Events: hipEvent_t handles for ua_run_circuit's per-gate timing.
StreamGraph: a HIP graph recorded by stream capture and replayed on a
stream, without the device-wide synchronisation torch's graph capture does.
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
        rt.hipStreamBeginCapture.argtypes = [ctypes.c_void_p, ctypes.c_int]
        rt.hipStreamBeginCapture.restype = ctypes.c_int
        rt.hipStreamEndCapture.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]
        rt.hipStreamEndCapture.restype = ctypes.c_int
        rt.hipGraphInstantiate.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]
        rt.hipGraphInstantiate.restype = ctypes.c_int
        rt.hipGraphLaunch.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        rt.hipGraphLaunch.restype = ctypes.c_int
        rt.hipGraphExecDestroy.argtypes = [ctypes.c_void_p]
        rt.hipGraphExecDestroy.restype = ctypes.c_int
        rt.hipGraphDestroy.argtypes = [ctypes.c_void_p]
        rt.hipGraphDestroy.restype = ctypes.c_int
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


_CAPTURE_RELAXED = 2   # hipStreamCaptureModeRelaxed


def _handle(stream) -> ctypes.c_void_p:
    return stream if isinstance(stream, ctypes.c_void_p) else ctypes.c_void_p(stream)


class StreamGraph:
    """Records fn() -- kernel launches on `stream` (a hipStream_t handle) and
    nothing that allocates, synchronises or sets memory -- into a HIP graph by
    stream capture; launch() replays it on a stream."""

    def __init__(self, stream, fn):
        rt = runtime()
        stream = _handle(stream)
        self._graph = ctypes.c_void_p()
        self._exec = ctypes.c_void_p()
        if rt.hipStreamBeginCapture(stream, _CAPTURE_RELAXED) != 0:
            raise RuntimeError("hipStreamBeginCapture failed")
        try:
            fn()
        finally:
            rc = rt.hipStreamEndCapture(stream, ctypes.byref(self._graph))
        if rc != 0:
            raise RuntimeError("hipStreamEndCapture failed")
        if rt.hipGraphInstantiate(ctypes.byref(self._exec), self._graph, None, None, 0) != 0:
            rt.hipGraphDestroy(self._graph)
            raise RuntimeError("hipGraphInstantiate failed")

    def launch(self, stream):
        if runtime().hipGraphLaunch(self._exec, _handle(stream)) != 0:
            raise RuntimeError("hipGraphLaunch failed")

    def __del__(self):
        if _rt is not None:
            if getattr(self, "_exec", None) and self._exec.value:
                _rt.hipGraphExecDestroy(self._exec)
            if getattr(self, "_graph", None) and self._graph.value:
                _rt.hipGraphDestroy(self._graph)
