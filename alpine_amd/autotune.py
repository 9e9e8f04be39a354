"""Launch-shape autotuner for the persistent traversal kernel.

Sweeps the `apg_trace_config` knobs -- workgroup size, persistent blocks per CU,
the idle-lane threshold at which a wave refills itself from the ray cursor,
and the LDS depth of the per-lane traversal stack -- on a caller-supplied ray
batch, timing each candidate with HIP events on the launch stream, and leaves
the fastest configuration set on the accelerator.
"""
from __future__ import annotations

import itertools
import time

BLOCK_SIZES = (128, 256)
BLOCKS_PER_CU = (2, 3, 4, 5, 6, 8)
FETCH_THRESHOLDS = (16, 24, 32, 40, 48, 64)
STACK_LDS = (8, 16)


def candidates(quick: bool = False):
    if quick:
        return [(256, b, f, s) for b, f, s in itertools.product((4, 5, 6, 8), (24, 32, 40), (8, 16))]
    return list(itertools.product(BLOCK_SIZES, BLOCKS_PER_CU, FETCH_THRESHOLDS, STACK_LDS))


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
    for bs, bpc, ft, st in candidates(quick):
        if bs * bpc > 256 * 16:
            continue
        cfg = type(base)()
        cfg.block_size, cfg.blocks_per_cu, cfg.fetch_threshold, cfg.stack_lds = bs, bpc, ft, st
        ms = time_config(acc, org, dir, cfg, reps, any_hit)
        rows.append((ms, {"block_size": bs, "blocks_per_cu": bpc, "fetch_threshold": ft,
                          "stack_lds": st}))
    rows.sort(key=lambda r: r[0])
    best = rows[0][1]
    cfg = type(base)()
    for k, v in best.items():
        setattr(cfg, k, v)
    acc.set_trace_config(cfg)
    if log:
        log(f"[autotune] {len(rows)} configs in {time.perf_counter() - t0:.1f} s; best {best} "
            f"{rows[0][0]:.3f} ms, default-like worst {rows[-1][0]:.3f} ms")
    return best, rows
