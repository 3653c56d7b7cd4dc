"""Independent device work of several variables, overlapped.

Each fuzzy variable's chain (grouping, all-pairs comparison, expansion,
sort) reads its size results back to the host a few times, which leaves the
GPU idle between launches when the variables run one after another. Here
every variable runs from its own host thread on its own stream (default
priority), so one variable's short kernels and readbacks overlap another's
comparison kernel. The library calls release the GIL (ctypes), and torch's
current stream is per thread. FASTER_SERIAL_VARS=1 runs the variables one
after another on the caller's stream (for per-kernel timing).
"""
from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor

import torch

from . import _lib

_streams = {}
# True: run() calls the tasks one after another on the caller's stream (the
# per-kernel timings of the work accounting then do not overlap).
SERIAL = False


def _side_streams(dev, k):
    pool = _streams.setdefault(dev.index, [])
    while len(pool) < k:
        pool.append(torch.cuda.Stream(device=dev))
    return pool[:k]


def run(tasks):
    """Runs the callables ``tasks`` (each returning a tensor or a tuple whose
    first element is a tensor) concurrently, each on a side stream ordered
    after the caller's current stream; returns their results in order once
    all of their device work has finished."""
    if len(tasks) <= 1 or SERIAL or os.environ.get("FASTER_SERIAL_VARS") == "1":
        return [t() for t in tasks]
    dev = _lib.device()
    main = torch.cuda.current_stream(dev)
    streams = _side_streams(dev, len(tasks))

    def call(fn, s):
        with torch.cuda.device(dev), torch.cuda.stream(s):
            s.wait_stream(main)
            return fn()

    with ThreadPoolExecutor(max_workers=len(tasks)) as ex:
        futures = [ex.submit(call, fn, s) for fn, s in zip(tasks, streams)]
        results = [f.result() for f in futures]
    for s in streams:
        main.wait_stream(s)
        s.synchronize()
    for r in results:
        t = r[0] if isinstance(r, tuple) else r
        if isinstance(t, torch.Tensor) and t.is_cuda:
            t.record_stream(main)
    return results


def record_on_current(*objs):
    """Marks every device tensor inside ``objs`` (tensors, lists, tuples,
    dataclass instances) as used on the current stream, for tensors that a
    side stream of run() allocated and the caller's stream keeps using."""
    import dataclasses

    main = torch.cuda.current_stream(_lib.device())

    def walk(o):
        if isinstance(o, torch.Tensor):
            if o.is_cuda:
                o.record_stream(main)
        elif isinstance(o, (list, tuple)):
            for x in o:
                walk(x)
        elif dataclasses.is_dataclass(o) and not isinstance(o, type):
            for f in dataclasses.fields(o):
                walk(getattr(o, f.name))

    for o in objs:
        walk(o)

