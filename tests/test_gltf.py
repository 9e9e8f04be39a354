"""The GLB reader of the `simple` app (alpine_amd.gltf), on the CPU.

Pinned parts: the float4x4 products of the node hierarchy and vertex transforms are bit-exact to
the reference's math/matrix.h (tests/golden/gltf_math.npz).  The rest of gltf_loader.cpp cannot be
built here (tinygltf / stb_image are absent), so the reader is checked against what its steps fix
in closed form: TRS matrices of known rotations, translation-only hierarchies that move vertices
exactly, the material mapping (metallicFactor > 0 -> Conductor, alpha = roughness), texel
conversion (d / 255, 2x - 1, the sRGB segments), PNG rows under every filter, sparse morph
targets, punctual-light positions and "weights" animations.
"""
import os

import numpy as np
import pytest

from alpine_amd import gltf, simple
from tests.glb_writer import GlbWriter, png_bytes

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "gltf_math.npz")


def test_matrix_products_match_reference():
    g = np.load(GOLD)
    for i in range(len(g["m0"])):
        np.testing.assert_array_equal(gltf.mat_mul(g["m0"][i], g["m1"][i]).view(np.uint32),
                                      g["prod"][i].view(np.uint32))
        np.testing.assert_array_equal(gltf.mat_apply(g["m0"][i], g["v"][i]).view(np.uint32),
                                      g["mv"][i].view(np.uint32))


def test_trs_matrix_closed_forms():
    np.testing.assert_array_equal(gltf.node_matrix({}), np.eye(4, dtype=np.float32).reshape(-1))
    s = np.float32(np.sqrt(0.5))
    m = gltf.node_matrix({"rotation": [0, 0, float(s), float(s)], "translation": [1, 2, 3],
                          "scale": [2, 3, 4]})
    # 90 degrees about z: x -> y, y -> -x (columns scaled by s), translation in column 3
    want = np.array([[0, -3, 0, 1], [2, 0, 0, 2], [0, 0, 4, 3], [0, 0, 0, 1]], np.float32)
    np.testing.assert_allclose(m.reshape(4, 4).T, want, atol=1e-6)
    mm = np.asarray(np.arange(16), np.float32)
    np.testing.assert_array_equal(gltf.node_matrix({"matrix": mm.tolist()}), mm)


def _rgb_image(h, w, seed):
    return np.random.default_rng(seed).integers(0, 256, (h, w, 3)).astype(np.uint8)


@pytest.mark.parametrize("chans", [3, 4])
def test_png_rows_under_every_filter(chans):
    img = np.random.default_rng(chans).integers(0, 256, (9, 7, chans)).astype(np.uint8)
    np.testing.assert_array_equal(gltf.decode_png(png_bytes(img)), img)


def test_png_palette_and_grey():
    pal = np.array([[10, 20, 30], [200, 100, 0], [5, 250, 128]], np.uint8)
    idx = np.random.default_rng(1).integers(0, 3, (5, 6)).astype(np.uint8)
    np.testing.assert_array_equal(gltf.decode_png(png_bytes(idx, palette=pal)), pal[idx])
    rgba = gltf.decode_png(png_bytes(idx, palette=pal, trns=[255, 0]))
    np.testing.assert_array_equal(rgba[..., 3], np.where(idx == 1, 0, 255))
    grey = gltf.decode_png(png_bytes(np.zeros((2, 2, 1), np.uint8)))
    assert grey.shape[2] == 1


def test_texel_conversion_closed_forms():
    d = np.array([[[0, 10, 11, 255]]], np.uint8)
    t = gltf.texture_texels(d[..., :3], signed=False, srgb=False)
    np.testing.assert_array_equal(t[0, 0], np.float32([0, 10, 11, 255]) / np.float32(255))
    s = gltf.texture_texels(d[..., :3], signed=True, srgb=False)
    np.testing.assert_array_equal(s[0, 0, :3], np.float32([0, 10, 11]) / np.float32(255) * np.float32(2) - np.float32(1))
    lin = gltf.texture_texels(np.array([[[0, 10, 255]]], np.uint8), signed=False, srgb=True)[0, 0]
    # 10/255 <= 0.04045: linear segment; 255/255 -> 1 exactly
    assert lin[0] == 0 and lin[1] == np.float32(np.float32(10) / np.float32(255) / np.float32(12.92)) and lin[2] == 1


def _scene_file(tmp_path):
    w = GlbWriter()
    tex = w.image(png_bytes(_rgb_image(4, 8, 3)))
    ntex = w.image(png_bytes(np.full((2, 2, 3), 128, np.uint8)))
    jpeg = len(w.gl.setdefault("images", []))
    w.gl["images"].append({"bufferView": w._view(b"\xff\xd8\xff\xe0not-a-png"), "mimeType": "image/jpeg",
                           "name": "photo"})
    w.gl["textures"].append({"source": jpeg})
    m_diel = w.material(base=(0.5, 0.25, 0.125, 1), metallic=0.0, rough=0.3, base_tex=tex, normal_tex=ntex)
    m_cond = w.material(base=(0.9, 0.8, 0.7, 1), rough=0.6)          # metallicFactor defaults to 1
    m_jpeg = w.material(metallic=0.0, base_tex=len(w.gl["textures"]) - 1)
    quad = dict(pos=[[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], idx=np.uint16([0, 1, 2, 0, 2, 3]),
                nrm=[[0, 0, 1]] * 4, uv=[[0, 0], [1, 0], [1, 1], [0, 1]],
                tan=[[1, 0, 0, 1]] * 4, material=m_diel,
                targets=[{"pos": ("sparse", [2], [[0, 0, 0.5]]), "nrm": [[0, 0, 0]] * 4},
                         {"pos": [[0.25, 0, 0]] * 4}])
    tri = dict(pos=[[0, 0, 0], [0, 2, 0], [0, 0, 2]], idx=np.uint32([0, 1, 2]), nrm=[[1, 0, 0]] * 3,
               material=m_cond)
    me0 = w.mesh([quad], weights=[0, 0])
    me1 = w.mesh([tri, dict(tri, material=m_jpeg)])
    child = w.node(root=False, mesh=me1, translation=[0, 0, 4])
    n0 = w.node(mesh=me0, translation=[10, 0, 0], children=[child])
    lt = w.point_light((1.0, 0.5, 0.25), 8.0)
    w.node(translation=[1, 2, 3], extensions={"KHR_lights_punctual": {"light": lt}})
    w.weights_animation(n0, [0.0, 1.0, 2.0], [[0.5, 0.0], [1.0, 0.5], [0.0, 1.0]])
    w.weights_animation(n0, [0.0, 1.0], [[0, 0], [1, 1]], interpolation="STEP")
    p = tmp_path / "scene.glb"
    p.write_bytes(w.bytes())
    return str(p)


def test_reader_follows_the_reference_loader(tmp_path):
    sc = gltf.read_glb(_scene_file(tmp_path))
    assert sorted(sc.meshes) == [0, 1]
    q, t = sc.meshes[0], sc.meshes[1]
    # translation-only hierarchy: positions move exactly, normals and tangents do not
    np.testing.assert_array_equal(q["verts"], np.float32([[10, 0, 0], [11, 0, 0], [11, 1, 0], [10, 1, 0]]))
    np.testing.assert_array_equal(t["verts"][:3], np.float32([[10, 0, 4], [10, 2, 4], [10, 0, 6]]))
    np.testing.assert_array_equal(q["normals"], np.float32([[0, 0, 1]] * 4))
    np.testing.assert_array_equal(q["tangents"], np.float32([[1, 0, 0, 1]] * 4))
    np.testing.assert_array_equal(q["tris"], [[0, 1, 2], [0, 2, 3]])
    # two primitives: indices offset by the first one's vertices, a material per triangle
    np.testing.assert_array_equal(t["tris"], [[0, 1, 2], [3, 4, 5]])
    np.testing.assert_array_equal(t["tri_mat"], [1, 2])
    # sparse and dense morph targets
    np.testing.assert_array_equal(q["targets"][0]["pos"], np.float32([[0, 0, 0], [0, 0, 0], [0, 0, 0.5], [0, 0, 0]]))
    np.testing.assert_array_equal(q["targets"][1]["pos"], np.float32([[0.25, 0, 0]] * 4))
    # materials: Dielectric for metallic 0, Conductor by default; alpha = (roughness, roughness)
    k0, bc0, a0, bt0, nt0 = sc.materials[0]
    k1, bc1, a1, bt1, nt1 = sc.materials[1]
    assert (k0, k1) == (0, 1) and a0 == (np.float32(0.3), np.float32(0.3)) and a1 == (np.float32(0.6),) * 2
    assert bc0 == (0.5, 0.25, 0.125) and bt1 is None and nt1 is None
    assert bt0 is not None and nt0 is not None
    np.testing.assert_array_equal(sc.textures[nt0][..., :3], np.float32(128) / np.float32(255) * 2 - 1)
    # the JPEG is reported and left out
    assert sc.materials[2][3] is None and any("photo" in m for m in sc.messages)
    # the point light sits at its node's origin with intensity colour x intensity
    (inten, pos), = sc.lights
    np.testing.assert_array_equal(pos, np.float32([1, 2, 3]))
    np.testing.assert_array_equal(inten, np.float32([8, 4, 2]))
    # one Animation per glTF animation; only LINEAR "weights" channels are kept
    assert len(sc.animations) == 2 and sc.animations[1] == []
    (mi, times, weights), = sc.animations[0]
    assert mi == 0 and times.tolist() == [0, 1, 2] and weights.shape == (3, 2)


def test_simple_reads_glb_before_touching_the_gpu(tmp_path, capsys):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    bad = tmp_path / "bad.glb"
    bad.write_bytes(b"glTF\x01\x00\x00\x00")
    assert simple.main(["-i", str(bad), "-o", str(tmp_path / "o.ppm"), "--spp", "1", "--accelerator", "gpu",
                        "--lightSampler", "bvh"]) == 1
    assert "ERROR: failed to load" in capsys.readouterr().out


def test_jpeg_textures_through_pillow(tmp_path):
    """A JPEG base-colour texture is decoded (Pillow / libjpeg) within JPEG's loss of the
    encoded pixels and converted like a PNG one; a grey JPEG is reported as the reference
    reports images of other channel counts."""
    PIL = pytest.importorskip("PIL.Image")
    import io
    rng = np.random.default_rng(4)
    img = np.repeat(np.repeat(rng.integers(0, 256, (4, 6, 3), dtype=np.uint8), 8, 0), 8, 1)

    def jpeg(a, mode):
        b = io.BytesIO()
        PIL.fromarray(a, mode).save(b, format="JPEG", quality=98, subsampling=0)
        return b.getvalue()

    w = GlbWriter()
    for data, name in ((jpeg(img, "RGB"), "colour"), (jpeg(img[..., 0], "L"), "grey")):
        w.gl.setdefault("images", []).append({"bufferView": w._view(data), "mimeType": "image/jpeg",
                                              "name": name})
        w.gl.setdefault("textures", []).append({"source": len(w.gl["images"]) - 1})
    m0 = w.material(metallic=0.0, base_tex=0)
    m1 = w.material(metallic=0.0, base_tex=1)
    tri = dict(pos=[[0, 0, 0], [1, 0, 0], [0, 1, 0]], idx=np.uint16([0, 1, 2]), material=m0)
    w.node(mesh=w.mesh([tri, dict(tri, material=m1)]))
    p = tmp_path / "jpeg.glb"
    p.write_bytes(w.bytes())
    sc = gltf.read_glb(str(p))
    t0 = sc.materials[0][3]
    assert t0 is not None and sc.materials[1][3] is None
    assert any("grey" in m and "channel count" in m for m in sc.messages)
    dec = gltf.decode_image(jpeg(img, "RGB"))
    assert dec.shape == img.shape and np.abs(dec.astype(int) - img).mean() < 3
    np.testing.assert_array_equal(sc.textures[t0], gltf.texture_texels(dec, signed=False, srgb=True))
