"""Launch-shape autotuner for the persistent traversal kernel.

Sweeps the `apg_trace_config` knobs -- workgroup size, persistent blocks per CU,
the idle-lane threshold at which a wave refills itself from the ray cursor,
the LDS depth of the per-lane traversal stack and the number of queued triangle
tests that starts a wave-wide triangle round -- on a caller-supplied ray batch,
timing each candidate with HIP events on the launch stream, and leaves the
fastest configuration set on the accelerator.
"""
from __future__ import annotations

import itertools
import time

BLOCK_SIZES = (128, 256)
BLOCKS_PER_CU = (2, 3, 4, 5, 6, 8)
FETCH_THRESHOLDS = (4, 8, 16, 24, 32, 48)
STACK_LDS = (8, 16)
TRI_BATCH = (32, 48, 64)
KNOBS = ("block_size", "blocks_per_cu", "fetch_threshold", "stack_lds", "tri_batch")


def candidates(quick: bool = False):
    if quick:
        return [(256, b, f, s, t) for b, f, s, t in
                itertools.product((4, 5, 6), (8, 16, 24), (8, 16), (32, 48, 64))]
    return list(itertools.product(BLOCK_SIZES, BLOCKS_PER_CU, FETCH_THRESHOLDS, STACK_LDS, TRI_BATCH))


def make_config(base, **knobs):
    """A copy of `base` with the named knobs replaced."""
    cfg = type(base)()
    for k in KNOBS:
        setattr(cfg, k, knobs.get(k, getattr(base, k)))
    return cfg


def time_config(acc, org, dir, cfg, reps=3, any_hit=False):
    import torch
    acc.set_trace_config(cfg)
    fn = (lambda: acc.intersect_any(org, dir)) if any_hit else (lambda: acc.intersect(org, dir))
    fn()
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def autotune(acc, org, dir, quick=False, reps=3, any_hit=False, log=None):
    """Return (best config dict, sorted [(ms, config dict)]); the best is left active."""
    t0 = time.perf_counter()
    base = acc.trace_config()
    rows = []
    for cand in candidates(quick):
        if cand[0] * cand[1] > 256 * 16:
            continue
        knobs = dict(zip(KNOBS, cand))
        ms = time_config(acc, org, dir, make_config(base, **knobs), reps, any_hit)
        rows.append((ms, knobs))
    rows.sort(key=lambda r: r[0])
    best = rows[0][1]
    acc.set_trace_config(make_config(base, **best))
    if log:
        log(f"[autotune] {len(rows)} configs in {time.perf_counter() - t0:.1f} s; best {best} "
            f"{rows[0][0]:.3f} ms, default-like worst {rows[-1][0]:.3f} ms")
    return best, rows
