"""HIP graphs of fml_amd calls (no reference counterpart).

A graph records every libfmlhip call the calling thread makes on one stream
between ``fmlh_graph_begin`` and ``fmlh_graph_end`` and replays the whole
sequence with one ``hipGraphLaunch``: e.g. the Coulomb matrices and the
ARAD kernels of a batch, whose tile list is planned on the device, so the
recorded work never waits on the host.

    call = AradKernelCall(q1, z1, n1, q2, z2, n2, sigmas)   # fml_amd.batch
    g = Graph.capture(call.run, keep=[call])
    g.launch()          # replays on torch's current stream

The captured calls read and write the device buffers they were given at
capture time, so only calls on buffers that outlive the graph may be
captured: the prepared calls of ``fml_amd.batch`` own their inputs, output
and workspace for that reason (the functions of ``fml_amd.kernels`` allocate
their results per call and are not for capture).  Buffer contents may change
between replays (new descriptors with the same atom counts, new coordinates
of the same molecules); sizes may not.  Capture runs on a side stream,
because the null stream cannot be captured.
"""
from __future__ import annotations

import ctypes

from . import _native


class Graph:
    """One instantiated HIP graph (an ``fmlh_graph``)."""

    def __init__(self):
        self._g = ctypes.c_void_p(0)
        self.keep = []   # objects whose device memory the graph reads

    @classmethod
    def capture(cls, fn, stream=None, keep=()):
        """Capture fn() on `stream` (a torch.cuda.Stream; default: a new side
        stream that waits on the current one)."""
        torch = _native.torch_mod()
        g = cls()
        g.keep = list(keep)
        cur = torch.cuda.current_stream()
        s = stream if stream is not None else torch.cuda.Stream()
        s.wait_stream(cur)
        with torch.cuda.stream(s):
            sp = ctypes.c_void_p(s.cuda_stream)
            _native.call("fmlh_graph_begin", sp)
            try:
                fn()
            except BaseException:
                bad = ctypes.c_void_p(0)
                _native.load().fmlh_graph_end(sp, ctypes.byref(bad))
                if bad.value:
                    _native.load().fmlh_graph_destroy(bad)
                raise
            _native.call("fmlh_graph_end", sp, ctypes.byref(g._g))
        cur.wait_stream(s)
        return g

    def launch(self, stream=None):
        """Replay on `stream` (default: torch's current stream)."""
        if not self._g.value:
            raise RuntimeError("graph was destroyed")
        torch = _native.torch_mod()
        s = stream if stream is not None else torch.cuda.current_stream()
        _native.call("fmlh_graph_launch", self._g, ctypes.c_void_p(s.cuda_stream))

    def destroy(self):
        if self._g.value:
            _native.call("fmlh_graph_destroy", self._g)
            self._g = ctypes.c_void_p(0)
        self.keep = []

    def __del__(self):
        try:
            if self._g.value and _native._lib is not None:
                _native._lib.fmlh_graph_destroy(self._g)
        except Exception:
            pass
