"""Morph-target animation feed (the reference's Animation, animation.cpp:10-39, and
Mesh::update, shapes/mesh.cpp:29-55) over the GPU accelerator.

``MorphAnimation.update(acc, time)`` clamps ``time`` to each channel's keyframe range, finds the
bracketing keyframes (the first whose successor's time is >= the clamped time), interpolates the
two weight vectors with C++20 ``std::lerp`` and blends the shape's vertices on the GPU with
``GpuAccelerator.morph`` (the shading normals and tangents follow the same weights).  The caller
rebuilds afterwards (``update_scene``), as the reference's updateScene does.
"""
from __future__ import annotations

import numpy as np

f32 = np.float32


def std_lerp(a, b, t):
    """C++20 std::lerp in float32, elementwise over the weight vectors.

    The published algorithm of P0811R3 ("Well-behaved interpolation for numbers and pointers",
    the one libstdc++ implements): exact at t = 0 and t = 1, monotonic, and t*b + (1-t)*a when
    a and b straddle zero."""
    a = np.asarray(a, f32)
    b = np.asarray(b, f32)
    t = f32(t)
    out = np.empty(np.broadcast(a, b).shape, f32)
    one = f32(1)
    for i, (x, y) in enumerate(zip(a.reshape(-1), b.reshape(-1))):
        if (x <= 0 and y >= 0) or (x >= 0 and y <= 0):
            r = f32(f32(t * y) + f32(f32(one - t) * x))
        elif t == one:
            r = y
        else:
            v = f32(x + f32(t * f32(y - x)))
            if (t > one) == (y > x):
                r = v if y < v else y
            else:
                r = v if y > v else y
        out.reshape(-1)[i] = r
    return out


def keyframe_weights(times, weights, time):
    """Animation::update for one channel: the morph weights at `time` (times ascending, weights
    (nkeys, ntargets))."""
    times = np.asarray(times, f32)
    weights = np.asarray(weights, f32)
    tc = f32(min(max(f32(time), times[0]), times[-1]))  # std::clamp(time, front, back)
    idx = len(times) - 2
    for i in range(len(times) - 1):
        if tc <= times[i + 1]:
            idx = i
            break
    t0, t1 = times[idx], times[idx + 1]
    t = f32(f32(tc - t0) / f32(t1 - t0))
    return std_lerp(weights[idx], weights[idx + 1], t)


class MorphAnimation:
    """Channels of (shape id, keyframe times, keyframe weights)."""

    def __init__(self, channels):
        self.channels = [(int(s), np.asarray(t, f32), np.asarray(w, f32)) for s, t, w in channels]

    def update(self, acc, time, stream=None):
        for shape, times, weights in self.channels:
            acc.morph(shape, keyframe_weights(times, weights, time), stream=stream)
