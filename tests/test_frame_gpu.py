"""Animated frames as one HIP graph (apg_frame_render): replayed frames against the same frames
run eagerly through the separate entry points (transform, GPU forest rebuild, reset, render),
and a replay whose rebuild is cut short on purpose (a host-side test hook; every kernel still
ends normally) reported by apg_frame_status and followed by a correct eager frame."""
import numpy as np
import pytest

import alpine_amd as aa
from alpine_amd import scenes

pytestmark = pytest.mark.gpu

APG_E_CAPACITY = -5
TILE, SPP, DEPTH, RR, LEAF = 32, 4, 5, 3, 2


def _setup(B):
    acc = aa.GpuAccelerator(0)
    for v, f in B.meshes:
        acc.append_mesh(v, f)
    acc.update_scene(leaf_threshold=LEAF, per_shape=True, stats=False)
    acc.set_materials([aa.material(*m) for m in B.materials], np.zeros(B.triangles, np.uint32))
    acc.set_lights([aa.light(**l) for l in B.lights], aa.LS_UNIFORM)
    n = len(B.meshes)
    acc.set_tiles([aa.camera(**c) for c in B.cams], TILE, TILE, roots=np.arange(n, dtype=np.uint32))
    acc.render_reset(B.tiles_x * TILE, B.tiles_y * TILE, 0)
    return acc


def _transforms(B, f):
    rot = B.rotation(0.2 + 0.05 * f)
    return [(k, first, count, rot) for k, (first, count) in enumerate(B.anim_range)]


def _eager_frame(acc, B, f, seed):
    for k, first, count, m in _transforms(B, f):
        acc.transform_vertices(k, m, first=first, count=count)
    acc.update_scene(leaf_threshold=LEAF, per_shape=True, stats=False)
    acc.render_reset(acc.width, acc.height, seed)
    acc.render(SPP, DEPTH, RR)
    rgb, _, total = acc.render_read()
    assert total == SPP
    return rgb


def _graph_frame(acc, B, f, seed):
    acc.frame_render(_transforms(B, f), SPP, DEPTH, RR, seed=seed, leaf_threshold=LEAF, per_shape=True)
    rgb, _, total = acc.render_read()
    assert total == SPP
    return rgb


def test_replayed_frames_equal_eager_frames():
    print("[gpu] frame graph: 8 scenes, 6 frames replayed vs eager", flush=True)
    B = scenes.BatchedScenes(n_scenes=8, lo=5_000, hi=80_000, seed=211, tile=TILE)
    g, e = _setup(B), _setup(B)
    for f in range(6):
        seed = 100 + 7 * f
        got = _graph_frame(g, B, f, seed)
        assert g.frame_status() == 0
        ref = _eager_frame(e, B, f, seed)
        print(f"frame {f}: counts {g.frame_counts()} mean {got.mean():.5f}", flush=True)
        assert np.isfinite(got).all() and got.mean() > 0.0
        # same trees (the wide layout is deterministic), same seeds: the same sums
        np.testing.assert_array_equal(got, ref)
    eager, captures, replays = g.frame_counts()
    assert eager == 1 and captures == 1 and replays == 5
    for k in range(len(B.meshes)):
        np.testing.assert_array_equal(np.array(g.get_vertex_buffer(k)), np.array(e.get_vertex_buffer(k)))


def test_short_replay_is_reported_and_next_frame_is_eager():
    print("[gpu] frame graph: replay with its rebuild cut short", flush=True)
    B = scenes.BatchedScenes(n_scenes=4, lo=5_000, hi=40_000, seed=223, tile=TILE)
    g, e = _setup(B), _setup(B)
    _graph_frame(g, B, 0, 5)                      # eager
    _eager_frame(e, B, 0, 5)
    g.frame_level_cut(1000)                       # the next capture runs no binary level
    _graph_frame(g, B, 1, 6)                      # captured, replayed: its rebuild stops at once
    assert g.frame_status() == APG_E_CAPACITY
    _eager_frame(e, B, 1, 6)                      # the reference applies frame 1's transforms too
    g.frame_level_cut(0)
    for f in (2, 3, 4):                           # eager, then captured again and replayed
        got = _graph_frame(g, B, f, 10 + f)
        assert g.frame_status() == 0
        np.testing.assert_array_equal(got, _eager_frame(e, B, f, 10 + f))
    assert g.frame_counts() == (2, 2, 3)


def test_frame_render_refuses_what_it_cannot_replay():
    import ctypes as C
    B = scenes.BatchedScenes(n_scenes=2, lo=2_000, hi=4_000, seed=5, tile=TILE)
    acc = _setup(B)
    lib = acc._lib
    rp = aa.RenderParams(SPP, DEPTH, RR, 0, aa._f3((0, 0, 0)), 1, 4, 1, 0.5, 0, 0, 0, 0, 0)
    assert lib.apg_frame_render(acc._h, None, 0, None, C.byref(rp), 0, None) == -1  # guide = 1
    rp.guide = 0
    assert lib.apg_frame_render(acc._h, None, 3, None, C.byref(rp), 0, None) == -1  # no transforms
    assert acc.frame_counts() == (0, 0, 0)
