"""A GLB scene through the `simple` app's loader onto the GPU, against the oracle integrator given
the same arrays: one mesh of two primitives (a textured, normal-mapped dielectric floor and a
conductor), normals, tangents, uvs, two morph targets animated to t = 0.5, a punctual light in the
file and the app's disk light, white background."""
import numpy as np
import pytest

import alpine_amd as aa
import oracle
from alpine_amd import gltf, scenes, simple
from alpine_amd.animation import keyframe_weights
from tests.glb_writer import GlbWriter, png_bytes

pytestmark = pytest.mark.gpu


def _glb(path):
    w = GlbWriter()
    rng = np.random.default_rng(4)
    tex = w.image(png_bytes(rng.integers(0, 256, (8, 8, 3)).astype(np.uint8)))
    nimg = np.full((4, 4, 3), 128, np.uint8)
    nimg[..., 0] = rng.integers(100, 160, (4, 4))
    nimg[..., 2] = 250
    ntex = w.image(png_bytes(nimg))
    m_floor = w.material(base=(0.8, 0.8, 0.8, 1), metallic=0.0, rough=0.4, base_tex=tex, normal_tex=ntex)
    m_metal = w.material(base=(0.9, 0.6, 0.3, 1), metallic=1.0, rough=0.3)
    q, g = scenes.quad_mesh((-2, -1, -2), (4, 0, 0), (0, 0, 4), 8, 8)
    b, bt = scenes.box_mesh(0.8, 0.8, 0.8, 25.0, (0.0, -1.0, 0.5), 3)
    nq, nb = len(q), len(b)
    floor = dict(pos=q, idx=np.uint32(g).reshape(-1), nrm=np.tile([0, 1, 0], (nq, 1)),
                 uv=q[:, [0, 2]] * 0.7, tan=np.tile([1, 0, 0, 1], (nq, 1)), material=m_floor,
                 targets=[{"pos": np.zeros((nq, 3)), "nrm": rng.normal(0, 0.2, (nq, 3)), "tan": np.zeros((nq, 3))},
                          {"pos": np.zeros((nq, 3)), "nrm": np.zeros((nq, 3)), "tan": rng.normal(0, 0.3, (nq, 3))}])
    bn = b - b.mean(0)
    bn /= np.linalg.norm(bn, axis=1, keepdims=True)
    box = dict(pos=b, idx=np.uint32(bt).reshape(-1), nrm=bn, uv=b[:, [0, 1]],
               tan=np.column_stack([np.cross(bn, [0, 1, 0.3]), np.ones(nb)]), material=m_metal,
               targets=[{"pos": np.tile([0, 0.3, 0], (nb, 1)), "nrm": np.zeros((nb, 3)), "tan": np.zeros((nb, 3))},
                        {"pos": rng.normal(0, 0.05, (nb, 3)), "nrm": np.zeros((nb, 3)), "tan": np.zeros((nb, 3))}])
    me = w.mesh([floor, box], weights=[0, 0])
    n = w.node(mesh=me, translation=[0.0, 0.2, 0.0])
    lt = w.point_light((1.0, 0.9, 0.7), 6.0)
    w.node(translation=[0.5, 1.5, -1.0], extensions={"KHR_lights_punctual": {"light": lt}})
    w.weights_animation(n, [0.0, 1.0], [[0.0, 1.0], [1.0, 0.0]])
    path.write_bytes(w.bytes())


def test_glb_scene_renders_like_the_oracle(tmp_path):
    print("[gpu] glb scene vs oracle", flush=True)
    p = tmp_path / "scene.glb"
    _glb(p)
    sc = gltf.read_glb(str(p))
    assert not sc.messages
    t_anim = 0.5
    W = H = 48
    spp, depth = 4, 4
    # GPU: the simple app's loading path
    acc = aa.GpuAccelerator(0)
    tri_mat, gmats, plights = simple.load_glb_scene(acc, sc, t_anim)
    acc.update_scene()
    acc.set_materials([aa.material(*m) for m in gmats], tri_mat)
    disk = dict(kind=aa.LIGHT_DISK, p0=(0.0, 10.0, 0.0), normal=(0.0, -1.0, 0.0), color=(1.0, 1.0, 1.0),
                power=100.0, radius=5.0)
    pl = [dict(kind=aa.LIGHT_POINT_INTENSITY, p0=pp, color=ii) for ii, pp in plights]
    acc.set_lights([aa.light(**disk)] + [aa.light(**l) for l in pl], aa.LS_BVH)
    cam = dict(eye=(0.0, 1.0, -3.5), target=(0.0, -0.6, 0.4), up=(0.0, 1.0, 0.0),
               fovy=float(np.pi / 2), aspect=1.0)
    acc.set_camera(aa.camera(**cam))
    acc.render_reset(W, H, 0)
    acc.render(spp, depth, 0, (1.0, 1.0, 1.0))
    rgb, _, total = acc.render_read()
    assert total == spp
    # oracle: the same mesh, morphed on the host (Mesh::update order), with the same attributes
    me = sc.meshes[0]
    (mi, times, kw), = sc.animations[0]
    wts = keyframe_weights(times, kw, t_anim)
    v = me["verts"].copy()
    for j, tg in enumerate(me["targets"]):
        v = (v + (tg["pos"] * wts[j]).astype(np.float32)).astype(np.float32)
    np.testing.assert_array_equal(acc.get_vertex_buffer(0).view(np.uint32), v.view(np.uint32))
    mats = [oracle.make_material(*m) for m in gmats]
    lights = [oracle.make_light(**disk)] + [oracle.make_light(**l) for l in pl]
    o = oracle.OracleScene(v, me["tris"], tri_mat, mats, lights, oracle.LS_BVH, oracle.make_camera(**cam),
                           (1.0, 1.0, 1.0))
    o.set_shading(me["normals"], None, me["tangents"], me["uvs"], None,
                  np.stack([t["nrm"] for t in me["targets"]]), np.stack([t["tan"] for t in me["targets"]]), wts)
    o.set_textures(sc.textures)
    _, ref = o.render(W, H, spp, depth, oracle.sampler_states(W * H, 0), None, 0, 8)
    ref = ref.astype(np.float32)
    close = np.abs(rgb - ref) <= 1e-3 * (1.0 + np.abs(ref))
    print("pixels equal within 1e-3:", close.mean(), "means", rgb.mean(), ref.mean())
    assert close.mean() > 0.97
    assert abs(rgb.mean() - ref.mean()) <= 0.01 * ref.mean()
