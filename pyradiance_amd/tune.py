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

DEFAULT = {"trace_kernel": 1, "refill": 16, "window": 1 << 24, "blocks_per_cu": 4}


def candidates() -> list:
    """The shapes tried by default: the one-ray-per-thread kernel, and the
    persistent kernel over refill thresholds, grid sizes and windows."""
    out = [dict(DEFAULT, trace_kernel=0)]
    for refill, bpc, win in itertools.product((16, 32, 48), (2, 4, 8), (1 << 22, 1 << 23, 1 << 24)):
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


def scene_candidates() -> list:
    """Tree-construction choices tried by `autotune_scene`: the builder (host
    SAH or device LBVH) and how many room-sized primitives stay out of the
    tree (tested first on every ray; 0 puts every primitive in the tree)."""
    return [{"builder": b, "max_big": m} for b in (_lib.BVH_SAH, _lib.BVH_LBVH) for m in (0, 8, 32)]


def autotune_scene(ctx: "_lib.Context", commit, run, cands: list | None = None) -> dict:
    """Time `run()` after `commit()` (which recommits the scene loaded on
    `ctx`; loading it again would add its objects a second time) under each
    tree-construction choice; recommit with the fastest.  Results do not
    depend on the tree; the walk's cost does.

    Returns {"best": choice, "ms": its time, "tried": [(choice, ms, commit_ms), ...]}."""
    cands = cands or scene_candidates()
    tried = []
    for cand in cands:
        ctx.set_bvh_builder(cand["builder"])
        ctx.set_tuning(max_big=cand["max_big"])
        t0 = time.perf_counter()
        commit()
        t_commit = time.perf_counter() - t0
        run()
        t0 = time.perf_counter()
        run()
        tried.append((cand, 1e3 * (time.perf_counter() - t0), 1e3 * t_commit))
    best, ms, _ = min(tried, key=lambda x: x[1])
    ctx.set_bvh_builder(best["builder"])
    ctx.set_tuning(max_big=best["max_big"])
    commit()
    return {"best": best, "ms": ms, "tried": tried}


def autotune_dcsky(A_ptr: int, B_ptr: int, C_ptr: int, M: int, K: int, N: int, reps: int = 10) -> dict:
    """Time the DC x sky product's tile forms on device buffers of the
    caller's own sizes (mode 2: operands loaded per wavefront, 32 x 64 tiles;
    mode 3: 128 x 128 tiles staged through LDS; mode 4, for K <= 160: rows of
    A held in registers, sky columns streamed through LDS) and return the
    fastest mode (the same product: all accumulate the same hi/lo split in
    FP32)."""
    import torch
    tried = []
    for mode in ((2, 3, 4) if K <= 160 else (2, 3)):
        _lib.dcsky_device(A_ptr, B_ptr, C_ptr, M, K, N, mode)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()           # the product runs on the null stream
        for _ in range(reps):
            _lib.dcsky_device(A_ptr, B_ptr, C_ptr, M, K, N, mode)
        e1.record()
        torch.cuda.synchronize()
        tried.append((mode, e0.elapsed_time(e1) / reps))
    best, ms = min(tried, key=lambda x: x[1])
    return {"best": best, "ms": ms, "tried": tried}
