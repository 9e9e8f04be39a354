"""GPU traversal with the watertight triangle test (apg_trace_config.triangle_test =
APG_TRI_WATERTIGHT), through the C ABI.

The kernels and apg_debug_triangles compile the same test (traverse_core.h) with the
same flags, so against the host brute force: hit t bit-exact, the same triangle except
exact-t ties, barycentrics bit-exact where the triangle agrees, any-hit equal to
"closest t < tmax".  Closed meshes seen from inside lose no ray, at the small size and
on the 1.3 M-triangle displaced sphere.
"""
import numpy as np
import pytest

import oracle
from alpine_amd import BUILDER_GPU, TRI_REFERENCE, TRI_WATERTIGHT, GpuAccelerator, scenes
from alpine_amd import debug_triangles
from tests.parity import MISS, classify
from tests.test_watertight import inside_rays_at_edges, inside_rays_at_vertices, unit_sphere

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, np.float32).T)).cuda()


def _host(hits):
    import torch
    torch.cuda.synchronize()
    p, t, b0, b1 = (x.cpu().numpy() for x in hits)
    return p.view(np.uint32), t, b0, b1


def _accel(v, f, lt=4):
    acc = GpuAccelerator(0)
    acc.append_mesh(v, f)
    acc.update_scene(leaf_threshold=lt, builder=BUILDER_GPU)
    acc.set_triangle_test(TRI_WATERTIGHT)
    assert acc.trace_config().triangle_test == TRI_WATERTIGHT
    return acc


def _check_against_brute(acc, v9, o, d, tmax=None):
    import torch
    g = _host(acc.intersect(_dev(o), _dev(d)))
    ref = debug_triangles(v9, o, d, watertight=True)
    c = classify(g, ref)
    assert c["unexplained"] == 0 and c["bary_mismatch"] == 0, c
    occ = acc.intersect_any(_dev(o), _dev(d)).cpu().numpy()
    np.testing.assert_array_equal(occ, ref[0] != MISS)
    if tmax is not None:
        occ = acc.intersect_any(_dev(o), _dev(d), torch.from_numpy(tmax).cuda()).cpu().numpy()
        np.testing.assert_array_equal(occ, (ref[0] != MISS) & (ref[1] < tmax))
    return g, c


@pytest.mark.parametrize("aim", ["vertices", "edges"])
def test_closed_sphere_no_leaks_matches_host(aim):
    print(f"[gpu] watertight closed sphere aimed at {aim}", flush=True)
    v, f = unit_sphere(3)
    v9 = v[f].reshape(-1, 9)
    if aim == "vertices":
        o, d = inside_rays_at_vertices(v, 20_000, seed=3)
    else:
        o, d = inside_rays_at_edges(v, f, 20_000, seed=4)
    acc = _accel(v, f)
    tmax = np.random.default_rng(6).uniform(0.2, 1.6, len(o)).astype(np.float32) * \
        np.linalg.norm(d, axis=1).astype(np.float32)
    g, c = _check_against_brute(acc, v9, o, d, tmax)
    assert (g[0] != MISS).all()
    print("classes", {k: v for k, v in c.items() if k != "unexplained_idx"})


def test_small_scene_random_rays_and_mode_switch(small_scene):
    """Watertight hits on the mixed scene against the host brute force; switching back
    to the reference test restores the oracle's hits bit for bit."""
    print("[gpu] watertight small scene + mode switch", flush=True)
    v, f = small_scene
    v9 = v[f].reshape(-1, 9)
    lo, hi = scenes.scene_bounds(v)
    o, d = scenes.incoherent_rays(4000, lo, hi, seed=21)
    acc = _accel(v, f)
    tmax = np.random.default_rng(22).uniform(0, np.linalg.norm(hi - lo), len(o)).astype(np.float32)
    g, c = _check_against_brute(acc, v9, o, d, tmax)
    assert c["hits"] > len(o) // 4
    acc.set_triangle_test(TRI_REFERENCE)
    ref = oracle.OracleBvh(v, f, 8).intersect(o, d)
    g = _host(acc.intersect(_dev(o), _dev(d)))
    c = classify(g, (ref.prim, ref.t, ref.b0, ref.b1))
    assert c["unexplained"] == 0 and c["bary_mismatch"] == 0, c


@pytest.mark.slow
def test_displaced_sphere_full_size_no_leaks():
    """1.3 M-triangle closed displaced sphere, 1 M rays from its centre aimed at its
    vertices: the watertight mode loses none (the reference test's losses are counted)."""
    print("[gpu] watertight displaced sphere 1M rays", flush=True)
    v, f = scenes.displaced_icosphere(8)
    rng = np.random.default_rng(31)
    n = 1 << 20
    o = np.zeros((n, 3), np.float32)
    d = v[rng.integers(0, len(v), n)].astype(np.float32)
    acc = _accel(v, f, 4)
    g = _host(acc.intersect(_dev(o), _dev(d)))
    assert (g[0] != MISS).all()
    assert acc.intersect_any(_dev(o), _dev(d)).cpu().numpy().all()
    acc.set_triangle_test(TRI_REFERENCE)
    m = _host(acc.intersect(_dev(o), _dev(d)))
    both = (m[0] != MISS)
    print("reference test misses", int((~both).sum()), "of", n)
    # where both tests hit, they agree to float rounding on t
    rel = np.abs(m[1][both].astype(np.float64) - g[1][both]) / g[1][both]
    assert np.max(rel) <= 2e-6
