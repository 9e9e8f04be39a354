"""Traversal stack bounds: the wide-depth guarantee, the overflow report, and builds of
degenerate input.

The reference's traversal keeps a 64-entry stack and only asserts against overflow
(bvh.cpp:218-231, bvh_common.h:15).  Here a lane pushes at most one entry per 8-wide level,
so a build deeper than the stack is refused (APG_E_CAPACITY) and a kernel that still runs out
(only possible under the stack_limit test hook) drops the push, flags the launch, and the
host reports APG_E_OVERFLOW.  The CPU tests run the kernel's per-lane loop on the host
(apg_debug_host_trace / _limited); the GPU test checks the error codes of the real kernel.
"""
import numpy as np
import pytest

from alpine_amd import ApgError, DebugScene, scenes

LEAF = 2


@pytest.fixture(scope="module")
def chain():
    v, f = scenes.geometric_chain()
    return v, f, DebugScene(v, f, LEAF)


def _centroid_rays(v, f, n, seed=3):
    """Rays from below the chain's first triangle towards random triangles' centroids."""
    rng = np.random.default_rng(seed)
    k = rng.integers(0, len(f), n)
    c = v[f[k]].astype(np.float64).mean(axis=1)
    org = np.stack([np.full(n, 0.5 * float(v[:, 0].min())), np.zeros(n), np.zeros(n)], -1)
    d = c - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return org.astype(np.float32), d.astype(np.float32)


def _per_ray_depth(ds, org, dir):
    return np.array([ds.trace(org[i:i + 1], dir[i:i + 1])[1] for i in range(len(org))])


def test_chain_is_deep(chain):
    v, f, ds = chain
    assert ds.wide_depth() >= 20, "the chain must need more than the 8 LDS stack entries"
    assert ds.info()["max_depth"] >= 28


def test_stack_never_exceeds_wide_depth(chain, small_scene):
    """One entry per level at most: the deepest stack is below the number of wide levels."""
    v, f, ds = chain
    org, d = scenes.chain_rays(3000, v)
    _, m = ds.trace(org, d)
    assert 8 < m <= ds.wide_depth() - 1
    sv, sf = small_scene
    s2 = DebugScene(sv, sf, LEAF)
    lo, hi = scenes.scene_bounds(sv)
    o2, d2 = scenes.incoherent_rays(3000, lo, hi)
    for any_hit in (False, True):
        _, m2 = s2.trace(o2, d2, any_hit=any_hit)
        assert m2 <= s2.wide_depth() - 1


def test_bounded_stack_flags_exactly_the_deep_rays(chain):
    """A ray drops a push iff its unbounded traversal would hold more than stack_limit
    entries; the rays that fit give bit-identical results."""
    v, f, ds = chain
    o1, d1 = scenes.chain_rays(150, v)
    o2, d2 = _centroid_rays(v, f, 150)
    org, d = np.concatenate([o1, o2]), np.concatenate([d1, d2])
    depth = _per_ray_depth(ds, org, d)
    (p0, t0, a0, b0), _ = ds.trace(org, d)
    assert (p0 != 0xFFFFFFFF).sum() > 50, "the centroid rays hit the chain"
    for limit in (4, 12, 19, int(depth.max())):
        over = np.array([ds.trace_limited(org[i:i + 1], d[i:i + 1], limit)[1] for i in range(len(org))])
        np.testing.assert_array_equal(over, (depth > limit).astype(over.dtype))
        (p, t, a, b), total = ds.trace_limited(org, d, limit)
        assert total == int((depth > limit).sum())
        ok = depth <= limit
        np.testing.assert_array_equal(p[ok], p0[ok])
        np.testing.assert_array_equal(t[ok].view(np.uint32), t0[ok].view(np.uint32))
        np.testing.assert_array_equal(a[ok].view(np.uint32), a0[ok].view(np.uint32))
        np.testing.assert_array_equal(b[ok].view(np.uint32), b0[ok].view(np.uint32))
    occ0, _ = ds.trace(org, d, any_hit=True)
    occ, total = ds.trace_limited(org, d, int(depth.max()), any_hit=True)
    assert total == 0
    np.testing.assert_array_equal(occ, occ0)


def test_builder_ends_when_surface_areas_overflow():
    """Coordinates up to 2^99: every candidate split's cost overflows to inf.  The
    reference's findSplit then keeps its default Split (min = max = 0), puts every primitive
    below and buildNode recurses without end (bvh_common.cpp:203-266, 302-313: the non-empty
    asserts compile out).  The node becomes a leaf here, and traversal still finds every hit
    brute force finds (the oracle's tree is the reference's and would not end either, so
    the brute force here is the kernels' own triangle test over every triangle)."""
    from alpine_amd import debug_triangles
    v, f = scenes.geometric_chain(n=100, ratio=2.0)
    ds = DebugScene(v, f, LEAF)
    _, _, ids, _ = ds.arrays()
    assert sorted(ids.tolist()) == list(range(len(f)))
    org, d = _centroid_rays(v, f, 200)
    (p, t, _, _), _ = ds.trace(org, d)
    bp, bt, _, _ = debug_triangles(v[f].reshape(-1, 9), org, d)
    assert (bp != 0xFFFFFFFF).sum() > 50
    np.testing.assert_array_equal(t.view(np.uint32), bt.view(np.uint32))
    assert np.all((p == bp) | (t == bt))


@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
def test_nonfinite_vertices_rejected_by_host_build(bad):
    v, f = scenes.geometric_chain(n=8)
    v = v.copy()
    v[5, 1] = bad
    with pytest.raises(ApgError):
        DebugScene(v, f, LEAF)


@pytest.mark.gpu
def test_overflow_is_reported_on_the_gpu(chain):
    """The kernel's bounded push under a lowered stack_limit: apg_trace_status and the per-ray
    host entry point return APG_E_OVERFLOW, the flag clears, and the default stack (which the
    chain needs 21 entries of, 13 of them in the HBM spill) traces the oracle's hits."""
    import torch

    import oracle
    from alpine_amd import GpuAccelerator

    v, f, ds = chain
    acc = GpuAccelerator(0)
    acc.append_mesh(v, f)
    acc.update_scene(leaf_threshold=LEAF)
    assert acc.stats()["wide_depth"] == ds.wide_depth(), "GPU and host builders: same tree depth"
    o1, d1 = scenes.chain_rays(4096, v)
    o2, d2 = _centroid_rays(v, f, 4096)
    org, dir = np.concatenate([o1, o2]), np.concatenate([d1, d2])
    o = torch.from_numpy(np.ascontiguousarray(org.T)).cuda()
    d = torch.from_numpy(np.ascontiguousarray(dir.T)).cuda()
    ref = oracle.OracleBvh(v, f, 8)
    h = ref.intersect(org, dir)
    occ_ref = ref.occluded(org, dir)

    def check():
        prim, t, _, _ = acc.intersect(o, d)
        occ = acc.intersect_any(o, d)
        acc.trace_status()
        gp = prim.cpu().numpy().view(np.uint32)
        gt = t.cpu().numpy()
        np.testing.assert_array_equal(gt.view(np.uint32), h.t.view(np.uint32))
        assert np.all((gp == h.prim) | (gt == h.t))
        np.testing.assert_array_equal(occ.cpu().numpy(), occ_ref)

    print("overflow test: default stack", flush=True)
    check()
    cfg = acc.trace_config()
    cfg.stack_limit = 12
    acc.set_trace_config(cfg)
    print("overflow test: stack_limit 12", flush=True)
    acc.intersect(o, d)
    with pytest.raises(ApgError) as e:
        acc.trace_status()
    assert e.value.code == -6
    acc.trace_status()  # reported once, then clear
    rays6 = np.ascontiguousarray(np.concatenate([org, dir], axis=1)[:4096])
    n = len(rays6)
    import ctypes as C
    from alpine_amd import _ptr
    prim = np.empty(n, np.uint32)
    t, b0, b1 = (np.empty(n, np.float32) for _ in range(3))
    occ = np.empty(n, np.uint8)
    host = lambda: acc._lib.apg_intersect_host(acc._h, _ptr(rays6), n, _ptr(prim), _ptr(t), _ptr(b0), _ptr(b1))
    assert host() == -6
    assert host() == -6, "each host batch reports its own overflow"
    assert acc._lib.apg_occluded_host(acc._h, _ptr(rays6), None, n, _ptr(occ)) == -6
    cfg.stack_limit = 0
    acc.set_trace_config(cfg)
    print("overflow test: default stack again", flush=True)
    check()
    assert host() == 0
    np.testing.assert_array_equal(t.view(np.uint32), h.t[:n].view(np.uint32))
    assert acc._lib.apg_occluded_host(acc._h, _ptr(rays6), None, n, _ptr(occ)) == 0
    np.testing.assert_array_equal(occ, occ_ref[:n])
    acc.close()


@pytest.mark.gpu
@pytest.mark.parametrize("bad", [np.nan, np.inf])
def test_nonfinite_vertices_rejected_by_append(bad):
    from alpine_amd import GpuAccelerator
    v, f = scenes.geometric_chain(n=8)
    acc = GpuAccelerator(0)
    v = v.copy()
    v[2, 0] = bad
    with pytest.raises(ApgError) as e:
        acc.append_mesh(v, f)
    assert e.value.code == -1
    good = acc.append_mesh(np.nan_to_num(v, posinf=1.0), f)
    buf = acc.get_vertex_buffer(good)
    buf[0, 0] = bad
    with pytest.raises(ApgError):
        acc.update_shape(good)
    acc.close()
