"""hipGraph capture of fixed libawig launch sequences.

A propagation loop that records observables every step issues about ten small launches per
step; at 256^2 the host then paces the GPU.  Such a loop is run in chunks whose launches
read every per-chunk input (factor tables, times, the output ring) from buffers at fixed
addresses, so one chunk can be captured once (torch.cuda.CUDAGraph over the current HIP
stream, which every libawig launch uses) and replayed for every later chunk after its
inputs are rewritten.  Captured sequences contain kernel launches only: no allocation,
memset or host synchronisation.

The Python cyclic collector is run before a capture and held off during it: a collection
inside the capture could finalise an unrelated dropped object whose finaliser frees device
memory (hipFree), which a capture in global mode does not allow (the capture is then
invalidated).  Plans dropped while a capture is open are destroyed after it (engine.Plan).

Set AWIG_GRAPHS=0 to run every chunk eagerly (same kernels, same order, same results).
"""
import gc
import os


def enabled():
    return os.environ.get("AWIG_GRAPHS", "1") != "0"


class Captured:
    """One captured launch sequence; ``replay()`` runs it on the current stream."""

    def __init__(self, fn):
        import torch
        from . import engine
        self.graph = torch.cuda.CUDAGraph()
        gc.collect()
        was_enabled = gc.isenabled()
        gc.disable()
        try:
            with torch.cuda.graph(self.graph):
                fn()
        finally:
            if was_enabled:
                gc.enable()
            engine.destroy_deferred_plans()

    def replay(self):
        self.graph.replay()


def run_chunk(cache, key, fn, full):
    """Run ``fn`` (a fixed launch sequence) through a cached graph when ``full`` (the chunk
    has the captured length) and graphs are enabled, else eagerly."""
    if not (full and enabled()):
        fn()
        return
    g = cache.get(key)
    if g is None:
        g = cache[key] = Captured(fn)
    g.replay()
