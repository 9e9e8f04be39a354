"""Launch-shape autotuner for the persistent traversal kernel and the render stages.

`autotune` sweeps the `apg_trace_config` knobs -- workgroup size, persistent blocks per CU,
the idle-lane threshold at which a wave refills itself from the ray cursor,
the LDS depth of the per-lane traversal stack and the number of queued triangle
tests that starts a wave-wide triangle round -- on a caller-supplied ray batch,
timing each candidate with HIP events on the launch stream, and leaves the
fastest configuration set on the accelerator.  `autotune_render` does the same for
a render: the paths in flight (sample slots per pixel) and the persistent grids of
the shade, shadow-visibility and resolve stages (`apg_render_launch`), timed over
whole passes of the caller's frame.
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


# Wavefront render stages: persistent 256-thread workgroups per CU of the shade, shadow-visibility
# and resolve stages (apg_render_launch), and the paths in flight (sample slots per pixel).
SHADE_BLOCKS = (4, 6, 8, 12, 16)
SHADOW_BLOCKS = (2, 4, 8)
RESOLVE_BLOCKS = (2, 4, 8)


def _time_render(acc, width, height, spp, max_depth, rr_depth, background, slots, reps, seed=9, **kw):
    import torch
    acc.render_reset(width, height, seed, slots=slots)
    acc.render(spp, max_depth, rr_depth, background, **kw)  # warm-up (and capture, with graph=True)
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        acc.render(spp, max_depth, rr_depth, background, **kw)
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def autotune_render(acc, width, height, spp, max_depth, rr_depth=0, background=(0.0, 0.0, 0.0),
                    slots=(1,), reps=2, log=None, **render_kw):
    """Tune the render launch on the frame the accelerator holds: first the sample slots per
    pixel (paths in flight = width x height x slots) at the default stage grids, then the shade,
    shadow and resolve grids one stage at a time at the best slots.  spp must be a multiple of
    every slot count.  Leaves the best grids set and returns (best dict, [(ms, dict)]) with
    samples per second per candidate; the caller resets the render before its own passes."""
    t0 = time.perf_counter()
    acc.set_render_launch()
    rows = []

    def run(sl, shade, shadow, resolve):
        acc.set_render_launch(shade, shadow, resolve)
        ms = _time_render(acc, width, height, spp, max_depth, rr_depth, background, sl, reps, **render_kw)
        cfg = {"slots": sl, "shade": shade, "shadow": shadow, "resolve": resolve}
        rows.append((ms, cfg))
        return ms

    d = acc.render_launch()
    best_sl = min(slots, key=lambda sl: run(sl, d["shade"], d["shadow"], d["resolve"]))
    best = {"shade": d["shade"], "shadow": d["shadow"], "resolve": d["resolve"]}
    for stage, values in (("shade", SHADE_BLOCKS), ("shadow", SHADOW_BLOCKS), ("resolve", RESOLVE_BLOCKS)):
        timed = []
        for v in values:
            trial = dict(best, **{stage: v})
            timed.append((run(best_sl, trial["shade"], trial["shadow"], trial["resolve"]), v))
        best[stage] = min(timed)[1]
    acc.set_render_launch(best["shade"], best["shadow"], best["resolve"])
    ms_best = min(ms for ms, c in rows if c == dict(best, slots=best_sl))
    out = dict(best, slots=best_sl, ms_per_call=round(ms_best, 4),
               samples_per_s=round(width * height * spp / (ms_best * 1e-3), 1))
    if log:
        log(f"[autotune] render: {len(rows)} launches in {time.perf_counter() - t0:.1f} s; best {out}")
    rows.sort(key=lambda r: r[0])
    return out, rows
