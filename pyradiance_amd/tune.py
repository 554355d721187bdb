"""Launch-shape autotuner for the wavefront renderer.

The renderer's results do not depend on its launch shape (rg_set_tuning:
traversal kernel, refill threshold of the persistent kernel, persistent grid
size, generation window); only the time does.  `autotune` times a list of
candidate shapes on the caller's own workload and leaves the fastest set on
the context.
"""

from __future__ import annotations

import itertools
import time

import numpy as np

from . import _lib

DEFAULT = {"trace_kernel": 1, "refill": 16, "window": 1 << 21, "blocks_per_cu": 4}


def candidates() -> list:
    """The shapes tried by default: the one-ray-per-thread kernel, and the
    persistent kernel over refill thresholds, grid sizes and windows."""
    out = [dict(DEFAULT, trace_kernel=0)]
    for refill, bpc, win in itertools.product((8, 16, 32), (2, 4, 8), (1 << 20, 1 << 21, 1 << 22)):
        out.append({"trace_kernel": 1, "refill": refill, "window": win, "blocks_per_cu": bpc})
    return out


def autotune(ctx: "_lib.Context", run, cands: list | None = None, reps: int = 1) -> dict:
    """Time `run()` (one call of the workload on `ctx`) under each candidate
    shape after one warm-up call; set and return the fastest.

    Returns {"best": shape, "ms": its time, "tried": [(shape, ms), ...]}."""
    cands = cands or candidates()
    tried = []
    for cand in cands:
        ctx.set_tuning(**cand)
        run()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            run()
            ts.append(time.perf_counter() - t0)
        tried.append((cand, 1e3 * float(np.min(ts))))
    best, ms = min(tried, key=lambda x: x[1])
    ctx.set_tuning(**best)
    return {"best": best, "ms": ms, "tried": tried}
