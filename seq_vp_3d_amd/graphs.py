"""HIP graph replay of the device-only segments of the solver.

A segment is a fixed sequence of libseqvp_hip.so calls that reads nothing back to the host until its end: the
trial (candidate map, certificate screen, spline warp, energies, one copy of the scalars to page-locked memory)
and the direction (gradient sweep, equilibrated DST direction, its dot products and step bound, one copy of the
scalars). Its graph is captured the first time a set of buffer addresses is seen and launched from then on, so an
inner iteration issues two graph launches instead of some thirty library calls. The accepted state, the trial and
the BB history change places by swapping buffers, so a pair cycles through a few address sets; each keeps its
graph (least recently used first out).

The null stream cannot be captured: capture runs on a private stream of the segment set (nothing executes while
capturing), and launches go to the caller's current stream. Capture only where no other host thread uses the device
at the same time: while a capture is open, the HIP runtime fails blocking copies and synchronisations of other
threads with hipErrorStreamCaptureUnsupported, even in thread-local capture mode (the multi-slot server therefore
issues its segments call by call).
"""
import ctypes
import threading
from collections import OrderedDict

import torch

from . import _lib

# Graphs of segment sets that were garbage-collected. They are destroyed at the next launch of any segment set
# rather than in the collector, which may run inside another thread's capture. Their pair's work has finished: a
# pair is dropped only after its last segment's scalars were read (a host wait on its stream).
_RETIRED = []
_RETIRED_LOCK = threading.Lock()


def _destroy_retired():
    with _RETIRED_LOCK:
        retired = _RETIRED[:]
        del _RETIRED[:]
    for exe in retired:
        _lib.raw("seqvp_graph_destroy", exe)


def key_of(*items):
    """Address key of the tensors (and plain values) a segment is captured with."""
    return tuple(t.data_ptr() if isinstance(t, torch.Tensor) else t for t in items)


class SegmentGraphs:
    """Captured graphs of one pair's segments, by (segment name, address key)."""

    def __init__(self, device, capacity=256):
        self.device = torch.device(device)
        self.capacity = int(capacity)
        self._graphs = OrderedDict()
        self._stream = None
        self.captures = 0
        self.launches = 0

    def run(self, key, body):
        """Launch the graph of `body` for `key` on the current stream, capturing it first if `key` is new. `body`
        enqueues library calls on torch's current stream and must not read results, synchronise or allocate."""
        if _RETIRED:
            _destroy_retired()
        exe = self._graphs.get(key)
        if exe is None:
            exe = self._capture(body)
            self._graphs[key] = exe
            if len(self._graphs) > self.capacity:
                _, old = self._graphs.popitem(last=False)
                torch.cuda.current_stream(self.device).synchronize()   # the evicted graph may still run
                _lib.call("seqvp_graph_destroy", old)
        else:
            self._graphs.move_to_end(key)
        _lib.call("seqvp_graph_launch", exe, _lib.stream_handle(self.device))
        self.launches += 1

    def _capture(self, body):
        if self._stream is None:
            self._stream = torch.cuda.Stream(device=self.device)
        exe = ctypes.c_void_p()
        with torch.cuda.stream(self._stream):
            handle = _lib.stream_handle(self.device)
            _lib.call("seqvp_graph_begin", handle)
            try:
                body()
            except BaseException:
                _lib.raw("seqvp_graph_end", handle, ctypes.byref(exe))
                _lib.raw("seqvp_graph_destroy", exe)
                raise
            _lib.call("seqvp_graph_end", handle, ctypes.byref(exe))
        self.captures += 1
        return exe

    def clear(self):
        if self._graphs:
            torch.cuda.current_stream(self.device).synchronize()
        while self._graphs:
            _, exe = self._graphs.popitem()
            _lib.raw("seqvp_graph_destroy", exe)

    def __del__(self):
        try:
            with _RETIRED_LOCK:
                _RETIRED.extend(self._graphs.values())
            self._graphs.clear()
        except Exception:       # interpreter shutdown: the process releases the graphs
            pass
