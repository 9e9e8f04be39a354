"""Shading attributes, textures, environment map, AOVs and morph targets on the GPU.

 - apg_mesh_morph blends the vertices on the GPU exactly as the reference's Mesh::update
   (tests/golden/shading_small.npz, written by the reference's own mesh.cpp / animation.cpp).
 - A Cornell room whose triangles carry per-vertex normals on their own index buffer, tangents,
   uvs, a base-colour texture, a normal map and morphed normal / tangent deltas, seen through an
   open side onto an environment map, renders like the oracle integrator with the same
   attributes: the first-hit albedo / normal AOVs agree bit for bit on almost every pixel
   (their arithmetic is +, *, / and sqrt only), the radiance within the tolerance of the
   plain-scene comparison (test_render_gpu.py), which covers the GPU's cosf/sinf/powf ulps.
"""
import os

import numpy as np
import pytest

import alpine_amd as aa
import oracle
from alpine_amd import scenes
from alpine_amd.animation import MorphAnimation, keyframe_weights

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "shading_small.npz")


def test_morph_matches_reference_mesh_update():
    print("[gpu] morph vs reference Mesh::update", flush=True)
    g = np.load(GOLD)
    acc = aa.GpuAccelerator(0)
    other = acc.append_mesh(np.float32([[5, 5, 5], [6, 5, 5], [5, 6, 5]]), [[0, 1, 2]])
    sid = acc.append_mesh(g["a_verts"], g["a_prims"])
    acc.set_targets(sid, g["a_tpos"])
    anim = MorphAnimation([(sid, g["a_key_times"], g["a_key_weights"])])
    for name in ("a", "b", "a"):
        anim.update(acc, float(g[name + "_time"]))
        v = acc.get_vertex_buffer(sid)
        np.testing.assert_array_equal(v.view(np.uint32), g[name + "_out_out_verts"].view(np.uint32))
    np.testing.assert_array_equal(acc.get_vertex_buffer(other), np.float32([[5, 5, 5], [6, 5, 5], [5, 6, 5]]))
    # the rebuilt tree traces the morphed sheet like the oracle BVH over the same vertices
    acc.update_scene()
    import torch
    verts = np.concatenate([np.float32([[5, 5, 5], [6, 5, 5], [5, 6, 5]]), g["a_out_out_verts"]])
    tris = np.concatenate([np.uint32([[0, 1, 2]]), g["a_prims"] + 3])
    rng = np.random.default_rng(5)
    n = 4096
    org = np.column_stack([rng.uniform(0, 1, n), rng.uniform(0, 1, n), np.full(n, 2.0)]).astype(np.float32)
    dir = np.tile(np.float32([0, 0, -1]), (n, 1))
    prim, t, _, _ = acc.intersect(torch.from_numpy(org.T.copy()).cuda(), torch.from_numpy(dir.T.copy()).cuda())
    ref = oracle.OracleBvh(verts, tris, 8).intersect(org, dir)
    gp = prim.cpu().numpy().view(np.uint32)
    np.testing.assert_array_equal(t.cpu().numpy().view(np.uint32), ref.t.view(np.uint32))
    assert np.all((gp == ref.prim) | (t.cpu().numpy() == ref.t))


def _textured_room(size):
    """The Cornell room with attributes on every triangle (flat per-triangle normals on their own
    index buffer, planar uvs, tangents), its front open onto an environment map."""
    sc = scenes.cornell_scene(size, size, spp=1, max_depth=4, wall_res=6, box_res=2)
    rng = np.random.default_rng(9)
    v, t = sc.verts.astype(np.float32), sc.tris.astype(np.uint32)
    e0, e1 = v[t[:, 1]] - v[t[:, 0]], v[t[:, 2]] - v[t[:, 0]]
    fn = np.cross(e0, e1)
    fn /= np.linalg.norm(fn, axis=1, keepdims=True)
    nn = 3 * len(t)
    normals = (np.repeat(fn, 3, axis=0) + rng.normal(0, 0.15, (nn, 3))).astype(np.float32)
    normal_idx = np.arange(nn, dtype=np.uint32).reshape(-1, 3)
    tang = np.cross(normals, np.float32([0.3, 1.0, 0.2]))
    tang /= np.linalg.norm(tang, axis=1, keepdims=True)
    tangents = np.column_stack([tang, np.where(rng.random(nn) < 0.5, -1.0, 1.0)]).astype(np.float32)
    uvs = np.column_stack([v[:, 0] * 1.7 + v[:, 1] * 0.3, v[:, 2] * 1.3 - v[:, 1]]).astype(np.float32)
    ck = (np.add.outer(np.arange(6), np.arange(10)) % 2).astype(np.float32)
    bc_tex = np.stack([0.2 + 0.6 * ck, 0.3 + 0.4 * ck, 0.5 - 0.3 * ck, np.ones_like(ck)], -1)
    nmap = rng.normal(0, 0.3, (4, 4, 4)).astype(np.float32)
    nmap[..., 2] = 1.0
    nmap[..., :3] /= np.linalg.norm(nmap[..., :3], axis=-1, keepdims=True)
    env = rng.uniform(0.0, 2.0, (16, 32, 3)).astype(np.float32)
    dn = rng.normal(0, 0.2, (2, nn, 3)).astype(np.float32)
    dt = rng.normal(0, 0.2, (2, nn, 3)).astype(np.float32)
    w = np.float32([0.7, -0.4])
    # grey walls textured + normal-mapped; red wall normal-mapped; green wall a rough conductor
    mats = [((0, (0.73, 0.73, 0.73), (0.4, 0.4)), 0, 1), ((0, (0.65, 0.05, 0.05), (1.0, 1.0)), None, 1),
            ((1, (0.12, 0.45, 0.15), (0.3, 0.3)), None, None)]
    # the open front: move the camera back so the walls' edges and the sky are both in view
    sc.camera = dict(sc.camera, eye=(1.8, 1.6, -3.0), target=(-0.2, 1.0, 0.3), fovy=float(np.radians(80.0)))
    return sc, dict(normals=normals, normal_idx=normal_idx, tangents=tangents, uvs=uvs,
                    textures=[bc_tex, nmap], env=env, dn=dn, dt=dt, w=w, mats=mats)


def _gpu(sc, at, spp, aovs):
    acc = aa.GpuAccelerator(0)
    sid = acc.append_mesh(sc.verts, sc.tris)
    acc.set_shading(sid, at["normals"], at["normal_idx"], at["tangents"], at["uvs"])
    acc.set_targets(sid, np.zeros((2, len(sc.verts), 3), np.float32), at["dn"], at["dt"])
    acc.morph(sid, at["w"])
    acc.update_scene()
    acc.set_textures(at["textures"])
    acc.set_materials([aa.material(*m, base_color_tex=b, normal_tex=n) for m, b, n in at["mats"]], sc.tri_mat)
    acc.set_lights([aa.light(**l) for l in sc.lights], sc.light_sampler)
    acc.set_camera(aa.camera(**sc.camera))
    acc.set_environment(at["env"])
    acc.render_reset(sc.width, sc.height, sc.seed)
    acc.render(spp, sc.max_depth, sc.rr_depth, sc.background, aovs=aovs)
    rgb, _, total = acc.render_read()
    alb, nrm = acc.render_read_aovs()
    return rgb, alb, nrm, total


def _oracle(sc, at, spp):
    mats = [oracle.make_material(*m, base_color_tex=b, normal_tex=n) for m, b, n in at["mats"]]
    s = oracle.OracleScene(sc.verts, sc.tris, sc.tri_mat, mats, [oracle.make_light(**l) for l in sc.lights],
                           sc.light_sampler, oracle.make_camera(**sc.camera), sc.background)
    s.set_shading(at["normals"], at["normal_idx"], at["tangents"], at["uvs"], None, at["dn"], at["dt"], at["w"])
    s.set_textures(at["textures"])
    s.set_environment(at["env"])
    aov = np.zeros((sc.height, sc.width, 6), np.float64)
    states = oracle.sampler_states(sc.width * sc.height, sc.seed)
    _, acc = s.render(sc.width, sc.height, spp, sc.max_depth, states, None, sc.rr_depth, 8, aov=aov)
    return acc.astype(np.float32), aov


def test_attributes_aovs_match_oracle():
    print("[gpu] textured room AOVs vs oracle", flush=True)
    sc, at = _textured_room(48)
    rgb, alb, nrm, total = _gpu(sc, at, 1, True)
    ref, aov = _oracle(sc, at, 1)
    assert total == 1
    ra, rn = aov[..., :3].astype(np.float32), aov[..., 3:].astype(np.float32)
    hit = np.any(rn != 0, axis=-1)
    assert 0.3 < hit.mean() < 0.97, "the view must show both the room and the environment"
    same_a = np.all(alb == ra, axis=-1)
    same_n = np.all(nrm == rn, axis=-1)
    print("AOV pixels identical: albedo", same_a.mean(), "normal", same_n.mean())
    assert same_a.mean() > 0.995 and same_n.mean() > 0.995
    # misses see the environment map: the radiance of a pixel whose camera ray leaves the room
    miss = ~hit
    close = np.abs(rgb - ref) <= 1e-3 * (1.0 + np.abs(ref))
    assert close[miss].mean() > 0.99


def test_attributes_radiance_matches_oracle():
    print("[gpu] textured room radiance vs oracle", flush=True)
    sc, at = _textured_room(48)
    rgb, _, _, _ = _gpu(sc, at, 4, False)
    ref, _ = _oracle(sc, at, 4)
    close = np.abs(rgb - ref) <= 1e-3 * (1.0 + np.abs(ref))
    print("pixels equal within 1e-3:", close.mean(), "means", rgb.mean(), ref.mean())
    assert close.mean() > 0.97
    assert abs(rgb.mean() - ref.mean()) <= 0.01 * ref.mean()


def test_aovs_sum_over_slots_and_clear_on_reset():
    print("[gpu] AOV slots and reset", flush=True)
    sc, at = _textured_room(24)
    acc = aa.GpuAccelerator(0)
    sid = acc.append_mesh(sc.verts, sc.tris)
    acc.set_shading(sid, at["normals"], at["normal_idx"], at["tangents"], at["uvs"])
    acc.update_scene()
    acc.set_textures(at["textures"])
    acc.set_materials([aa.material(*m, base_color_tex=b, normal_tex=n) for m, b, n in at["mats"]], sc.tri_mat)
    acc.set_lights([aa.light(**l) for l in sc.lights], sc.light_sampler)
    acc.set_camera(aa.camera(**sc.camera))
    acc.render_reset(sc.width, sc.height, 3, slots=4)
    acc.render(8, 3, aovs=True)
    alb8, nrm8 = acc.render_read_aovs()
    acc.render_reset(sc.width, sc.height, 3, slots=4)
    acc.render(4, 3, aovs=True)
    alb4, _ = acc.render_read_aovs()
    hit = np.any(nrm8 != 0, axis=-1)
    assert hit.mean() > 0.5
    # the reset cleared the sums: the 4-sample render is the first pass of the 8-sample one, whose
    # second pass only adds base colours (>= 0) of the samples that hit
    assert np.all(alb8 >= alb4) and alb8.sum() > 1.5 * alb4.sum()
    # textures: the grey walls' base colour is the checker, never the material's own 0.73 grey
    assert not np.any(np.isclose(alb4[hit] / 4.0, 0.73).all(-1))
