"""The drop-in adapter (alpine_amd/adapter: GpuAccelerator / GpuLightSampler) through alpine's own
Accelerator and LightSampler interfaces, beside the reference's own Bvh and light samplers.

oracle/_ref/libadapter_check.so (oracle/adapter_check.cpp, built by oracle/Makefile from the
reference's unmodified sources) makes the same appendMesh / updateScene / intersect /
intersectAny calls on both accelerators -- per ray, from 8 threads for the GPU one, as alpine's
tile tasks would -- and the same LightSampler::sample calls on both samplers.
"""
import numpy as np
import pytest

import oracle
from alpine_amd import scenes

needs_lib = pytest.mark.skipif(not oracle.adapter_available(),
                               reason="oracle/_ref/libadapter_check.so is built only where the reference tree is")


def _ptr(a):
    return a.ctypes.data


@needs_lib
def test_adapter_without_gpu_behaves_like_the_dummy_accelerators():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    # INVALID_SHAPE_ID, nullptr, empty optional, false, empty optional (wide_bvh_dummy.cpp:9-53)
    assert oracle.adapter_lib().adapter_check_no_device() == 31


def test_dummy_adapter_returns_the_interface_sentinels():
    """gpu_dummy.cpp (the build without ALPINE_ENABLE_HIP) compiled against alpine's headers:
    every call returns the sentinel alpine's own dummy accelerators return, GPU or not."""
    import ctypes
    import os
    path = os.path.join(os.path.dirname(oracle.__file__), "_ref", "libdummy_check.so")
    if not os.path.exists(path):
        pytest.skip("dummy adapter not built (the reference tree is needed to compile it)")
    assert ctypes.CDLL(path).adapter_check_no_device() == 31


@needs_lib
@pytest.mark.gpu
def test_adapter_per_ray_queries_match_reference_bvh():
    print("[gpu] adapter per-ray vs reference Bvh", flush=True)
    verts, tris = scenes.small_scene(n_random=1500, subdiv=3)
    verts = np.ascontiguousarray(verts, np.float32)
    tris = np.ascontiguousarray(tris, np.uint32)
    nt = len(tris)
    starts = np.uint32([0, nt // 3, (2 * nt) // 3, nt])
    lo, hi = scenes.scene_bounds(verts)
    org, dir = scenes.incoherent_rays(3000, lo, hi, seed=21)
    co, cd = scenes.coherent_rays(32, 32)
    rays = np.ascontiguousarray(np.concatenate([np.hstack([org, dir]), np.hstack([co, cd])]), np.float32)
    n = len(rays)
    tfar = np.random.default_rng(3).uniform(0, np.linalg.norm(hi - lo), n).astype(np.float32)
    out = {}
    for side in "gr":
        out[side] = dict(shape=np.empty(n, np.uint32), prim=np.empty(n, np.uint32), t=np.empty(n, np.float32),
                         b=np.empty((n, 2), np.float32), ng=np.empty((n, 3), np.float32), occ=np.empty(n, np.uint8))
    g, r = out["g"], out["r"]
    rc = oracle.adapter_lib().adapter_check_trace(
        _ptr(verts), len(verts), _ptr(tris), _ptr(starts), 3, _ptr(rays), _ptr(tfar), n, 8,
        *[_ptr(g[k]) for k in ("shape", "prim", "t", "b", "ng", "occ")],
        *[_ptr(r[k]) for k in ("shape", "prim", "t", "b", "ng", "occ")])
    assert rc == 0
    hit = r["prim"] != 0xFFFFFFFF
    assert 0.2 < hit.mean() < 0.95
    np.testing.assert_array_equal(g["t"].view(np.uint32), r["t"].view(np.uint32))
    same = (g["shape"] == r["shape"]) & (g["prim"] == r["prim"])
    # hit ids equal except exact-t ties, where the first-found triangle may differ
    assert np.all(same | (g["t"] == r["t"]))
    assert same.mean() > 0.999
    np.testing.assert_array_equal(g["b"][same].view(np.uint32), r["b"][same].view(np.uint32))
    np.testing.assert_array_equal(g["ng"][same].view(np.uint32), r["ng"][same].view(np.uint32))
    np.testing.assert_array_equal(g["occ"], r["occ"])


def _lights(n, seed):
    rng = np.random.default_rng(seed)
    spec = np.zeros((n, 12), np.float32)
    spec[:, 0] = (rng.random(n) < 0.3).astype(np.float32)   # 30 % point lights
    spec[:, 1:4] = rng.uniform(-5, 5, (n, 3))
    nrm = rng.standard_normal((n, 3))
    spec[:, 4:7] = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    spec[:, 7:10] = rng.uniform(0.2, 1.0, (n, 3))
    spec[:, 10] = np.exp(rng.normal(0, 1.5, n))
    spec[:, 11] = rng.uniform(0.05, 0.5, n)
    return spec


@needs_lib
@pytest.mark.gpu
@pytest.mark.parametrize("kind", [0, 1, 2])
def test_adapter_light_sampler_matches_reference(kind):
    print(f"[gpu] adapter light sampler kind={kind}", flush=True)
    spec = _lights(37, 5 + kind)
    rng = np.random.default_rng(17)
    n = 3000
    u = rng.random(n).astype(np.float32)
    hit = rng.uniform(-6, 6, (n, 3)).astype(np.float32)
    ns = rng.standard_normal((n, 3)).astype(np.float32)
    ns /= np.linalg.norm(ns, axis=1, keepdims=True)
    gi, ri = np.empty(n, np.int32), np.empty(n, np.int32)
    gp, rp = np.empty(n, np.float32), np.empty(n, np.float32)
    rc = oracle.adapter_lib().adapter_check_light_sampler(kind, _ptr(spec), len(spec), _ptr(u), _ptr(hit),
                                                          _ptr(ns), n, _ptr(gi), _ptr(gp), _ptr(ri), _ptr(rp))
    assert rc == 0
    # the shading core is bit-exact to these TUs on the host; on the GPU sqrtf and divisions are
    # correctly rounded, so the choices and pdfs agree except at rare ulp-level ties
    assert (gi == ri).mean() > 0.999
    same = gi == ri
    np.testing.assert_allclose(gp[same], rp[same], rtol=1e-6)
