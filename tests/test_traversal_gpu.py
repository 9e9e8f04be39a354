"""GPU parity of the wide-BVH traversal kernels against the oracle (through the C ABI).

Gate: closest-hit IDs and any-hit booleans bit-exact except exact-t ties
(reported), hit t bit-exact (tolerance 1e-6 relative is the spec; the kernel
runs the reference's arithmetic so equality is expected), barycentrics
bit-exact where the triangle agrees.
"""
import numpy as np
import pytest

import oracle
from alpine_amd import BUILDER_GPU, BUILDER_HOST, GpuAccelerator, scenes
from tests.parity import MISS, classify, max_rel_t_err

pytestmark = pytest.mark.gpu

T_REL_TOL = 1e-6


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, np.float32).T)).cuda()


def _host(hits):
    import torch
    torch.cuda.synchronize()
    p, t, b0, b1 = (x.cpu().numpy() for x in hits)
    return p.view(np.uint32), t, b0, b1


def _accel(verts, tris, builder, lt=4):
    acc = GpuAccelerator(0)
    acc.append_mesh(verts, tris)
    acc.update_scene(leaf_threshold=lt, builder=builder)
    return acc


def _builders():
    return [BUILDER_HOST, BUILDER_GPU]


@pytest.mark.parametrize("lt,key,sah", [(8, "leaf_keys8", "sah8"), (16, "leaf_keys16", "sah16")])
def test_gpu_builder_reproduces_reference_tree(golden, lt, key, sah):
    """GPU binned SAH with the reference's leaf threshold: same leaf sets and SAH cost."""
    print(f"[gpu] gpu builder tree lt={lt}", flush=True)
    acc = _accel(golden["verts"], golden["tris"], BUILDER_GPU, lt)
    np.testing.assert_array_equal(acc.leaf_keys(), golden[key])
    assert acc.stats()["sah_cost"] == pytest.approx(float(golden[sah]), rel=1e-12)
    tri9, ids = acc.triangles()
    assert np.array_equal(np.sort(ids), np.arange(len(golden["tris"])))


@pytest.mark.parametrize("builder", _builders())
def test_golden_closest_and_any(golden, builder):
    print(f"[gpu] golden closest/any builder={builder}", flush=True)
    acc = _accel(golden["verts"], golden["tris"], builder)
    org, dir = _dev(golden["org"]), _dev(golden["dir"])
    g = _host(acc.intersect(org, dir))
    c = classify(g, (golden["prim"], golden["t"], golden["b0"], golden["b1"]))
    assert c["unexplained"] == 0 and c["bary_mismatch"] == 0, c
    assert c["agree"] + c["tie"] == c["n"]
    hit = golden["prim"] != MISS
    assert max_rel_t_err(g[1], golden["t"], hit) <= T_REL_TOL
    occ = acc.intersect_any(org, dir).cpu().numpy()
    np.testing.assert_array_equal(occ, golden["occ"])
    import torch
    tf = torch.from_numpy(golden["tfar"]).cuda()
    occ = acc.intersect_any(org, dir, tf).cpu().numpy()
    np.testing.assert_array_equal(occ, golden["occ_tfar"])


@pytest.mark.parametrize("builder", _builders())
def test_small_scene_random_rays(small_scene, builder):
    print(f"[gpu] small scene random rays builder={builder}", flush=True)
    v, f = small_scene
    lo, hi = scenes.scene_bounds(v)
    o, d = scenes.incoherent_rays(50_000, lo, hi, seed=41)
    ob = oracle.OracleBvh(v, f, 8)
    ref = ob.intersect(o, d, nthreads=8)
    acc = _accel(v, f, builder)
    g = _host(acc.intersect(_dev(o), _dev(d)))
    c = classify(g, (ref.prim, ref.t, ref.b0, ref.b1))
    assert c["unexplained"] == 0 and c["bary_mismatch"] == 0, c
    occ = acc.intersect_any(_dev(o), _dev(d)).cpu().numpy()
    np.testing.assert_array_equal(occ, ob.occluded(o, d, nthreads=8))


@pytest.mark.parametrize("builder", _builders())
def test_empty_inputs_and_scene(builder):
    import torch
    print(f"[gpu] empty inputs builder={builder}", flush=True)
    acc = GpuAccelerator(0)
    acc.update_scene(builder=builder)  # no meshes: every ray misses
    o = _dev(np.zeros((5, 3)))
    d = _dev(np.tile([0, 0, 1.0], (5, 1)))
    p, t, b0, b1 = _host(acc.intersect(o, d))
    assert np.all(p == MISS) and np.all(t == np.float32(3.4028235e38))
    assert acc.intersect_any(o, d).cpu().numpy().sum() == 0
    # n = 0 is a no-op
    e = torch.empty((3, 0), dtype=torch.float32, device="cuda")
    acc.intersect(e, e)


def test_multi_mesh_offsets_and_axis_rays(small_scene):
    print("[gpu] multi mesh + axis-aligned rays", flush=True)
    v, f = small_scene
    half = len(f) // 2
    acc = GpuAccelerator(0)
    s0 = acc.append_mesh(v, f[:half])
    s1 = acc.append_mesh(v, f[half:])
    assert (s0, s1) == (0, 1) and acc.prim_offset(1) == half
    acc.update_scene(builder=BUILDER_GPU)
    lo, hi = scenes.scene_bounds(v)
    rng = np.random.default_rng(5)
    o = rng.uniform(lo, hi, (4000, 3)).astype(np.float32)
    axes = np.concatenate([np.eye(3), -np.eye(3)]).astype(np.float32)
    d = axes[rng.integers(0, 6, 4000)]
    ref = oracle.OracleBvh(v, f, 8).intersect(o, d)
    g = _host(acc.intersect(_dev(o), _dev(d)))
    c = classify(g, (ref.prim, ref.t, ref.b0, ref.b1))
    assert c["unexplained"] == 0, c


@pytest.mark.parametrize("builder", _builders())
def test_vertex_update_rebuild(small_scene, builder):
    """getVertexBuffer / updateShape / updateScene: a moved mesh is traced where it moved to."""
    print(f"[gpu] vertex update + rebuild builder={builder}", flush=True)
    v, f = small_scene
    acc = GpuAccelerator(0)
    sid = acc.append_mesh(v, f)
    acc.update_scene(builder=builder)
    vb = acc.get_vertex_buffer(sid)
    vb[:, 0] += np.float32(0.5)
    acc.update_shape(sid)
    acc.update_scene(builder=builder)
    moved = v.copy()
    moved[:, 0] += np.float32(0.5)
    lo, hi = scenes.scene_bounds(moved)
    o, d = scenes.incoherent_rays(20_000, lo, hi, seed=3)
    ref = oracle.OracleBvh(moved, f, 8).intersect(o, d)
    g = _host(acc.intersect(_dev(o), _dev(d)))
    c = classify(g, (ref.prim, ref.t, ref.b0, ref.b1))
    assert c["unexplained"] == 0, c


def test_counted_run_same_results(small_scene):
    print("[gpu] counted run", flush=True)
    v, f = small_scene
    lo, hi = scenes.scene_bounds(v)
    o, d = scenes.incoherent_rays(30_000, lo, hi, seed=7)
    acc = _accel(v, f, BUILDER_GPU)
    g = _host(acc.intersect(_dev(o), _dev(d)))
    out, nodes, tris = acc.intersect_counted(_dev(o), _dev(d))
    g2 = _host(out)
    for a, b in zip(g, g2):
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    assert nodes >= len(o) and tris > 0


@pytest.mark.slow
def test_microbench_scene_full_ray_sets(microbench_scene):
    """1.51 M triangles; 4 M incoherent + 1 M coherent rays, all checked against the oracle."""
    import torch
    print("[gpu] microbench full ray sets", flush=True)
    v, f = microbench_scene
    lo, hi = scenes.scene_bounds(v)
    ob = oracle.OracleBvh(v, f, 8)
    acc8 = _accel(v, f, BUILDER_GPU, 8)
    np.testing.assert_array_equal(acc8.leaf_keys(), ob.leaf_keys())
    ratio = acc8.stats()["sah_cost"] / ob.stats()["sah_cost"]
    print("gpu-built / reference SAH cost", ratio, "build ms", acc8.stats()["build_ms"])
    assert ratio <= 1.02
    del acc8
    acc = _accel(v, f, BUILDER_GPU)
    o, d = scenes.incoherent_rays(4_194_304, lo, hi, seed=8)
    g = _host(acc.intersect(_dev(o), _dev(d)))
    ref = ob.intersect(o, d, nthreads=16)
    c = classify(g, (ref.prim, ref.t, ref.b0, ref.b1))
    print("incoherent closest", {k: v for k, v in c.items() if k != "unexplained_idx"})
    assert c["unexplained"] == 0 and c["bary_mismatch"] == 0, c
    occ = acc.intersect_any(_dev(o), _dev(d)).cpu().numpy()
    np.testing.assert_array_equal(occ, ob.occluded(o, d, nthreads=16))
    tf = np.random.default_rng(9).uniform(0, np.linalg.norm(hi - lo), len(o)).astype(np.float32)
    occ = acc.intersect_any(_dev(o), _dev(d), torch.from_numpy(tf).cuda()).cpu().numpy()
    np.testing.assert_array_equal(occ, ob.occluded(o, d, tf, nthreads=16))
    oc, dc = scenes.coherent_rays(1024, 1024)
    g = _host(acc.intersect(_dev(oc), _dev(dc)))
    ref = ob.intersect(oc, dc, nthreads=16)
    c = classify(g, (ref.prim, ref.t, ref.b0, ref.b1))
    print("coherent closest", {k: v for k, v in c.items() if k != "unexplained_idx"})
    assert c["unexplained"] == 0 and c["bary_mismatch"] == 0, c


def test_per_shape_forest_traces_each_ray_in_its_own_tree():
    """Independent scenes that overlap in space, built as one forest: ray i sees only the
    shape roots[i] names, exactly as the oracle does on that shape alone."""
    import torch
    print("[gpu] per-shape forest", flush=True)
    meshes = [scenes.random_triangles(n, seed, extent=1.5, lmin=0.1, lmax=0.5)
              for n, seed in ((3000, 11), (200, 12), (5000, 13))]
    acc = GpuAccelerator(0)
    for v, f in meshes:
        acc.append_mesh(v, f)
    acc.update_scene(per_shape=True)
    assert acc.stats()["wide_nodes"] >= 3
    rng = np.random.default_rng(4)
    n = 6000
    org = rng.uniform(-3, 3, (n, 3)).astype(np.float32)
    d = rng.standard_normal((n, 3))
    dir = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    roots = rng.integers(0, 3, n).astype(np.uint32)
    got = _host(acc.intersect_roots(_dev(org), _dev(dir),
                                    torch.from_numpy(roots.view(np.int32)).cuda()))
    occ = acc.intersect_any_roots(_dev(org), _dev(dir),
                                  torch.from_numpy(roots.view(np.int32)).cuda()).cpu().numpy()
    for s, (v, f) in enumerate(meshes):
        sel = roots == s
        ob = oracle.OracleBvh(v, f, 8)
        h = ob.intersect(org[sel], dir[sel])
        off = acc.prim_offset(s)
        ref_p = np.where(h.prim == MISS, MISS, h.prim + off).astype(np.uint32)
        c = classify(tuple(x[sel] for x in got), (ref_p, h.t, h.b0, h.b1))
        print("shape", s, c)
        assert c["unexplained"] == 0 and c["agree"] + c["tie"] == c["n"]
        np.testing.assert_array_equal(occ[sel], ob.occluded(org[sel], dir[sel]))


def test_animated_forest_rebuild_matches_oracle():
    """GPU rigid transform of a vertex range, then a per-shape rebuild without host round trip:
    hits equal the oracle's on the transformed geometry read back from the library."""
    import torch
    print("[gpu] animated forest rebuild", flush=True)
    from alpine_amd.scenes import BatchedScenes
    meshes = [scenes.random_triangles(2000, 21, extent=1.5, lmin=0.1, lmax=0.5),
              scenes.random_triangles(4000, 22, extent=1.5, lmin=0.1, lmax=0.5)]
    acc = GpuAccelerator(0)
    for v, f in meshes:
        acc.append_mesh(v, f)
    acc.update_scene(per_shape=True, stats=False)
    nv1 = len(meshes[1][0])
    for frame in range(3):
        acc.transform_vertices(1, BatchedScenes.rotation(0.3), first=nv1 // 2, count=nv1 - nv1 // 2)
        acc.update_scene(per_shape=True, stats=False)
    rng = np.random.default_rng(8)
    n = 4000
    org = rng.uniform(-2, 2, (n, 3)).astype(np.float32)
    d = rng.standard_normal((n, 3))
    dir = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    roots = np.full(n, 1, np.uint32)
    got = _host(acc.intersect_roots(_dev(org), _dev(dir), torch.from_numpy(roots.view(np.int32)).cuda()))
    v1 = np.array(acc.get_vertex_buffer(1))
    assert not np.array_equal(v1, meshes[1][0]) and np.array_equal(v1[: nv1 // 2], meshes[1][0][: nv1 // 2])
    h = oracle.OracleBvh(v1, meshes[1][1], 8).intersect(org, dir)
    off = acc.prim_offset(1)
    ref_p = np.where(h.prim == MISS, MISS, h.prim + off).astype(np.uint32)
    c = classify(got, (ref_p, h.t, h.b0, h.b1))
    print(c)
    assert c["unexplained"] == 0 and c["agree"] + c["tie"] == c["n"] and c["hits"] > 100
