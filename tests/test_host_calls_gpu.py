"""Host-memory per-ray entry points (apg_intersect_host / apg_occluded_host) called from many
threads at once: concurrent calls are coalesced into shared launches, and every call gets
exactly the results of the batched traversal for its own rays."""
import numpy as np
import pytest
import torch

import alpine_amd as aa
from alpine_amd import scenes

pytestmark = pytest.mark.gpu


def _scene():
    rng = np.random.default_rng(5)
    c = rng.uniform(-1.0, 1.0, (20000, 1, 3))
    v = (c + rng.normal(0.0, 0.02, (20000, 3, 3))).reshape(-1, 3).astype(np.float32)
    f = np.arange(len(v), dtype=np.uint32).reshape(-1, 3)
    return v, f


def _setup():
    v, f = _scene()
    acc = aa.GpuAccelerator(0)
    acc.append_mesh(v, f)
    acc.update_scene()
    lo, hi = scenes.scene_bounds(v)
    o, d = scenes.incoherent_rays(12347, lo, hi, seed=3)
    return acc, o, d


def _batched(acc, o, d):
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a.T)).cuda()
    prim, t, _, _ = (x.cpu().numpy() for x in acc.intersect(dev(o), dev(d)))
    occ = acc.intersect_any(dev(o), dev(d)).cpu().numpy()
    return prim.view(np.uint32), t, occ


@pytest.mark.parametrize("threads,per_call", [(64, 1), (256, 1), (17, 5), (3, 1000)])
def test_concurrent_host_calls_match_the_batched_traversal(threads, per_call):
    print(f"[gpu] host calls {threads} threads x {per_call} rays per call", flush=True)
    acc, o, d = _setup()
    prim, t, occ = _batched(acc, o, d)
    rays6 = np.concatenate([o, d], 1)
    (hp, ht), sec = acc.host_rate(rays6, threads, per_call)
    np.testing.assert_array_equal(hp, prim)
    np.testing.assert_array_equal(ht.view(np.uint32), t.view(np.uint32))
    hocc, _ = acc.host_rate(rays6, threads, per_call, any_hit=True)
    np.testing.assert_array_equal(hocc, occ)
    print(f"  {len(o) / sec / 1e6:.3f} Mrays/s closest hit", flush=True)
    assert (prim != 0xFFFFFFFF).sum() > 1000
