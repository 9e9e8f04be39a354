"""GLB scenes for the `simple` app: the reference's glTF path (loaders/gltf_loader.cpp) restated over
this module's own GLB / PNG readers (the reference uses tinygltf and stb_image, absent offline).

What is read, as the reference reads it:
  * every scene's root nodes are traversed with the node matrices multiplied down the hierarchy
    (matrix, or T * R * S from translation / rotation / scale; gltf_loader.cpp:122-203); a node's
    mesh becomes shape `mesh index` (the last node naming a mesh wins), its KHR_lights_punctual
    light a point light of intensity colour x intensity at the node's origin (:460-472);
  * a mesh concatenates its triangle primitives (u16 / u32 indices offset by the vertices before
    them; every triangle gets its primitive's material) with POSITION, NORMAL, TEXCOORD_0 and
    TANGENT as floats, and morph targets' POSITION (dense or sparse), NORMAL and TANGENT
    (:205-273, 275-373); positions are transformed as points, normals and tangents as
    directions by the node matrix (not renormalised, :253-270);
  * a material is a Conductor when metallicFactor > 0, else a Dielectric, with
    alpha = (roughnessFactor, roughnessFactor), the base colour factor's rgb, the base-colour
    texture (sRGB decoded to linear) and the normal texture (mapped to [-1, 1]) (:384-460);
  * animations keep the "weights" channels of LINEAR samplers on nodes with meshes (:474-530).
Images: PNG (8-bit RGB / RGBA / palette) is decoded here; baseline and progressive JPEG through
Pillow (libjpeg) where it is importable (stb_image's own JPEG decoder, which the reference uses,
may differ from libjpeg by a few code values: its IDCT and chroma upsampling are its own); other
images (16-bit, grey, JPEG without Pillow) are reported and left out, as the reference does for
images of other channel counts.
"""
from __future__ import annotations

import json
import struct
import zlib

import numpy as np

f32 = np.float32
_COMPONENT = {5120: np.int8, 5121: np.uint8, 5122: np.int16, 5123: np.uint16, 5125: np.uint32,
              5126: np.float32}
_NCOMP = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4, "MAT2": 4, "MAT3": 9, "MAT4": 16}


class GltfError(ValueError):
    pass


# ---- float4x4 in the reference's column-major layout (math/matrix.h:65-127) ----
def mat_identity():
    return np.eye(4, dtype=f32).reshape(-1)  # symmetric, so the layout does not matter


def mat_column(m, i):
    return m[4 * i: 4 * i + 4]


def mat_apply(m, v):
    """mul(float4x4, float4): c0 * v0 + c1 * v1 + c2 * v2 + c3 * v3, left to right."""
    v = np.asarray(v, f32)
    r = (mat_column(m, 0) * v[..., 0:1]).astype(f32)
    for i in (1, 2, 3):
        r = (r + (mat_column(m, i) * v[..., i:i + 1]).astype(f32)).astype(f32)
    return r


def mat_mul(a, b):
    """mul(float4x4, float4x4): column i of the product is mul(a, column i of b)."""
    out = np.empty(16, f32)
    for i in range(4):
        out[4 * i: 4 * i + 4] = mat_apply(a, mat_column(b, i))
    return out


def node_matrix(node):
    """The node's local matrix (getMatrix, gltf_loader.cpp:122-188)."""
    if "matrix" in node:
        return np.asarray(node["matrix"], np.float64).astype(f32)
    q = [f32(x) for x in node.get("rotation", (0.0, 0.0, 0.0, 1.0))]
    s = [f32(x) for x in node.get("scale", (1.0, 1.0, 1.0))]
    t = [f32(x) for x in node.get("translation", (0.0, 0.0, 0.0))]
    x, y, z, w = q
    one, two = f32(1), f32(2)
    m = np.zeros(16, f32)
    m[0] = (one - two * y * y - two * z * z) * s[0]
    m[1] = (two * x * y + two * z * w) * s[0]
    m[2] = (two * x * z - two * y * w) * s[0]
    m[4] = (two * x * y - two * z * w) * s[1]
    m[5] = (one - two * x * x - two * z * z) * s[1]
    m[6] = (two * y * z + two * x * w) * s[1]
    m[8] = (two * x * z + two * y * w) * s[2]
    m[9] = (two * y * z - two * x * w) * s[2]
    m[10] = (one - two * x * x - two * y * y) * s[2]
    m[12], m[13], m[14], m[15] = t[0], t[1], t[2], one
    return m


# ---- PNG (8-bit, non-interlaced) ----
def _unfilter(raw, h, stride, bpp):
    out = np.zeros((h, stride), np.uint8)
    prior = np.zeros(stride, np.int32)
    pos = 0
    for r in range(h):
        ftype = raw[pos]
        line = np.frombuffer(raw, np.uint8, stride, pos + 1).astype(np.int32)
        pos += stride + 1
        if ftype == 0:
            cur = line
        elif ftype == 2:
            cur = (line + prior) & 255
        elif ftype == 1:
            cur = line.copy()
            for i in range(bpp, stride):
                cur[i] = (cur[i] + cur[i - bpp]) & 255
        elif ftype == 3:
            cur = line.copy()
            for i in range(stride):
                left = cur[i - bpp] if i >= bpp else 0
                cur[i] = (cur[i] + ((left + prior[i]) >> 1)) & 255
        elif ftype == 4:
            cur = line.copy()
            for i in range(stride):
                a = int(cur[i - bpp]) if i >= bpp else 0
                b = int(prior[i])
                c = int(prior[i - bpp]) if i >= bpp else 0
                p = a + b - c
                pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
                pred = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
                cur[i] = (cur[i] + pred) & 255
        else:
            raise GltfError(f"bad PNG filter {ftype}")
        out[r] = cur
        prior = cur.astype(np.int32)
    return out


def decode_png(data):
    """(height, width, channels) uint8 for 8-bit RGB / RGBA / palette PNGs; channels as stb_image
    reports them (palette: 3, or 4 with transparency).  Raises GltfError otherwise."""
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise GltfError("not a PNG")
    pos, idat, plte, trns, ihdr = 8, [], None, None, None
    while pos < len(data):
        n, kind = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8: pos + 8 + n]
        pos += 12 + n
        if kind == b"IHDR":
            ihdr = struct.unpack(">IIBBBBB", body)
        elif kind == b"PLTE":
            plte = np.frombuffer(body, np.uint8).reshape(-1, 3)
        elif kind == b"tRNS":
            trns = np.frombuffer(body, np.uint8)
        elif kind == b"IDAT":
            idat.append(body)
        elif kind == b"IEND":
            break
    if ihdr is None:
        raise GltfError("PNG without IHDR")
    w, h, depth, ctype, _, _, interlace = ihdr
    if depth != 8 or interlace != 0 or ctype not in (0, 2, 3, 4, 6):
        raise GltfError(f"PNG bit depth {depth} / colour type {ctype} / interlace {interlace} not read")
    chans = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    rows = _unfilter(zlib.decompress(b"".join(idat)), h, w * chans, chans)
    img = rows.reshape(h, w, chans)
    if ctype == 3:
        if plte is None:
            raise GltfError("palette PNG without PLTE")
        idx = img[..., 0]
        rgb = plte[idx]
        if trns is not None:
            a = np.full(len(plte), 255, np.uint8)
            a[:len(trns)] = trns
            return np.concatenate([rgb, a[idx][..., None]], -1)
        return rgb
    return img


def decode_image(data):
    """(height, width, channels) uint8 of an embedded PNG or JPEG image."""
    if data[:8] == b"\x89PNG\r\n\x1a\n":
        return decode_png(data)
    if data[:3] == b"\xff\xd8\xff":
        try:
            import io
            from PIL import Image
        except ImportError:
            raise GltfError("JPEG images need Pillow") from None
        try:
            with Image.open(io.BytesIO(data)) as im:
                im.load()
                if im.mode not in ("L", "RGB"):
                    im = im.convert("RGB")   # CMYK / YCCK as stb_image returns them: 3 channels
                a = np.asarray(im, np.uint8)
        except (OSError, ValueError, SyntaxError) as e:
            raise GltfError(f"JPEG not decoded ({e})") from None
        return a[..., None] if a.ndim == 2 else a
    raise GltfError("unknown image format")


def texture_texels(img, signed, srgb):
    """createTexture (gltf_loader.cpp:417-458): float4 texels d / 255 (alpha 1 for RGB), then
    x * 2 - 1 for normal maps, then sRGB -> linear on rgb for base colours."""
    h, w, c = img.shape
    t = np.empty((h, w, 4), f32)
    t[..., :3] = img[..., :3].astype(f32) / f32(255)
    t[..., 3] = img[..., 3].astype(f32) / f32(255) if c == 4 else f32(1)
    if signed:
        t = (t * f32(2) - f32(1)).astype(f32)
    if srgb:
        c3 = t[..., :3]
        lin = np.where(c3 <= f32(0.04045), c3 / f32(12.92),
                       np.power((c3 + f32(0.055)) / f32(1.055), f32(2.4)))
        t[..., :3] = lin.astype(f32)
    return t


class GltfScene:
    def __init__(self):
        self.meshes = {}      # mesh index -> dict(verts, tris, tri_mat, normals, tangents, uvs, targets, weights)
        self.materials = []   # (kind, base colour, alpha, base-colour texture, normal texture)
        self.textures = []    # float4 texel arrays (h, w, 4)
        self.lights = []      # (intensity rgb, position xyz) point lights
        self.animations = []  # lists of channels (mesh index, key times, key weights (nkeys, ntargets))
        self.messages = []


def read_glb(path):
    with open(path, "rb") as f:
        data = f.read()
    if len(data) < 20 or data[:4] != b"glTF":
        raise GltfError("not a GLB file")
    version, length = struct.unpack("<II", data[4:12])
    if version != 2:
        raise GltfError(f"glTF version {version}")
    pos, gl, bin_chunk = 12, None, b""
    while pos + 8 <= min(length, len(data)):
        n, kind = struct.unpack("<II", data[pos:pos + 8])
        body = data[pos + 8: pos + 8 + n]
        if kind == 0x4E4F534A:
            gl = json.loads(body.decode("utf-8"))
        elif kind == 0x004E4942:
            bin_chunk = body
        pos += 8 + n
    if gl is None:
        raise GltfError("GLB without a JSON chunk")
    return _Loader(gl, bin_chunk).load()


class _Loader:
    def __init__(self, gl, bin_chunk):
        self.gl = gl
        self.bin = bin_chunk
        self.out = GltfScene()
        self._tex_cache = {}

    def _view(self, vi):
        v = self.gl["bufferViews"][vi]
        if v.get("buffer", 0) != 0 or self.gl["buffers"][v.get("buffer", 0)].get("uri"):
            raise GltfError("only the GLB's own binary chunk is read")
        return v

    def accessor(self, ai):
        a = self.gl["accessors"][ai]
        dt = np.dtype(_COMPONENT[a["componentType"]]).newbyteorder("<")
        nc = _NCOMP[a["type"]]
        count = a["count"]
        if "bufferView" in a:
            v = self._view(a["bufferView"])
            off = v.get("byteOffset", 0) + a.get("byteOffset", 0)
            stride = v.get("byteStride", 0) or nc * dt.itemsize
            if stride == nc * dt.itemsize:
                arr = np.frombuffer(self.bin, dt, count * nc, off).reshape(count, nc)
            else:
                arr = np.stack([np.frombuffer(self.bin, dt, nc, off + i * stride) for i in range(count)])
            arr = arr.copy()
        else:
            arr = np.zeros((count, nc), dt)
        return arr

    def _sparse_positions(self, ai, count):
        """A morph target's POSITION deltas, dense or sparse (gltf_loader.cpp:300-337)."""
        a = self.gl["accessors"][ai]
        sp = a.get("sparse")
        if not sp:
            return self.accessor(ai).astype(f32)
        out = np.zeros((count, 3), f32)
        iv = self._view(sp["indices"]["bufferView"])
        it = np.dtype(_COMPONENT[sp["indices"]["componentType"]]).newbyteorder("<")
        idx = np.frombuffer(self.bin, it, sp["count"], iv.get("byteOffset", 0) + sp["indices"].get("byteOffset", 0))
        vv = self._view(sp["values"]["bufferView"])
        vals = np.frombuffer(self.bin, "<f4", 3 * sp["count"],
                             vv.get("byteOffset", 0) + sp["values"].get("byteOffset", 0)).reshape(-1, 3)
        out[idx.astype(np.int64)] = vals
        return out

    def load(self):
        gl = self.gl
        for mi in range(len(gl.get("materials", []))):
            self.out.materials.append(self.material(mi))
        for sc in gl.get("scenes", []):
            for ni in sc.get("nodes", []):
                self.traverse(mat_identity(), ni)
        for anim in gl.get("animations", []):
            self.out.animations.append(self.animation(anim))
        return self.out

    def traverse(self, parent, ni):
        node = self.gl["nodes"][ni]
        m = mat_mul(parent, node_matrix(node))
        if "mesh" in node:
            self.out.meshes[node["mesh"]] = self.mesh(m, node["mesh"])
        lt = node.get("extensions", {}).get("KHR_lights_punctual")
        if lt is not None:
            src = self.gl["extensions"]["KHR_lights_punctual"]["lights"][lt["light"]]
            col = np.asarray(src.get("color", (1.0, 1.0, 1.0)), np.float64).astype(f32)
            intensity = (col * f32(src.get("intensity", 1.0))).astype(f32)
            pos = mat_apply(m, np.array([0, 0, 0, 1], f32))[:3]
            self.out.lights.append((intensity, pos))
        for c in node.get("children", []):
            self.traverse(m, c)

    def mesh(self, m, mi):
        src = self.gl["meshes"][mi]
        verts, normals, uvs, tangents, tris, tri_mat = [], [], [], [], [], []
        targets = None
        nv = 0
        for p in src["primitives"]:
            if p.get("mode", 4) != 4:
                raise GltfError("only triangle primitives are read")
            if "indices" not in p or "material" not in p:
                raise GltfError("primitives need indices and a material")
            idx = self.accessor(p["indices"]).reshape(-1)
            if idx.dtype not in (np.uint16, np.uint32, np.dtype("<u2"), np.dtype("<u4")):
                raise GltfError("indices must be u16 or u32")
            t = idx[: len(idx) // 3 * 3].astype(np.uint32).reshape(-1, 3) + np.uint32(nv)
            tris.append(t)
            tri_mat.append(np.full(len(t), p["material"], np.uint32))
            at = p["attributes"]
            pos = self.accessor(at["POSITION"]).astype(f32)
            verts.append(pos)
            for name, dst in (("NORMAL", normals), ("TEXCOORD_0", uvs), ("TANGENT", tangents)):
                if name in at:
                    dst.append(self.accessor(at[name]).astype(f32))
            if p.get("targets"):
                if targets is None:
                    targets = [dict(pos=[], nrm=[], tan=[]) for _ in p["targets"]]
                for tg, st in zip(targets, p["targets"]):
                    if "POSITION" in st:
                        tg["pos"].append(self._sparse_positions(st["POSITION"], len(pos)))
                    if "NORMAL" in st:
                        tg["nrm"].append(self.accessor(st["NORMAL"]).astype(f32))
                    if "TANGENT" in st:
                        tg["tan"].append(self.accessor(st["TANGENT"]).astype(f32)[:, :3])
            nv += len(pos)
        v = np.concatenate(verts)
        out = dict(verts=None, tris=np.concatenate(tris), tri_mat=np.concatenate(tri_mat), normals=None,
                   tangents=None, uvs=None, targets=None, weights=len(src.get("weights", [])))
        ones = np.ones((len(v), 1), f32)
        out["verts"] = mat_apply(m, np.hstack([v, ones]))[:, :3].copy()
        cat = lambda xs: np.concatenate(xs) if xs and sum(len(x) for x in xs) == len(v) else None
        n = cat(normals)
        if n is not None:
            out["normals"] = mat_apply(m, np.hstack([n, 0 * ones]))[:, :3].copy()
        tg = cat(tangents)
        if tg is not None and n is not None:
            txyz = mat_apply(m, np.hstack([tg[:, :3], 0 * ones]))[:, :3]
            out["tangents"] = np.hstack([txyz, tg[:, 3:4]]).astype(f32)
        out["uvs"] = cat(uvs)
        if targets:
            out["targets"] = [dict(pos=cat(t["pos"]), nrm=cat(t["nrm"]), tan=cat(t["tan"])) for t in targets]
        return out

    def texture(self, ti, signed, srgb):
        key = (ti, signed, srgb)
        if key in self._tex_cache:
            return self._tex_cache[key]
        img_i = self.gl["textures"][ti]["source"]
        im = self.gl["images"][img_i]
        name = im.get("name", str(img_i))
        result = None
        try:
            if "bufferView" not in im:
                raise GltfError("external image")
            v = self._view(im["bufferView"])
            data = self.bin[v.get("byteOffset", 0): v.get("byteOffset", 0) + v["byteLength"]]
            img = decode_image(bytes(data))
            if img.shape[2] in (3, 4):
                self.out.textures.append(texture_texels(img, signed, srgb))
                result = len(self.out.textures) - 1
            else:
                self.out.messages.append(f"The channel count of {name} is not valid")
        except (GltfError, zlib.error) as e:
            self.out.messages.append(f"image {name} not read: {e}")
        self._tex_cache[key] = result
        return result

    def material(self, mi):
        src = self.gl["materials"][mi]
        pbr = src.get("pbrMetallicRoughness", {})
        bc = np.asarray(pbr.get("baseColorFactor", (1.0, 1.0, 1.0, 1.0))[:3], np.float64).astype(f32)
        rough = f32(pbr.get("roughnessFactor", 1.0))
        bct = self.texture(pbr["baseColorTexture"]["index"], False, True) if "baseColorTexture" in pbr else None
        nt = self.texture(src["normalTexture"]["index"], True, False) if "normalTexture" in src else None
        kind = 1 if pbr.get("metallicFactor", 1.0) > 0.0 else 0
        return (kind, tuple(bc.tolist()), (float(rough), float(rough)), bct, nt)

    def animation(self, anim):
        chans = []
        for ch in anim.get("channels", []):
            tgt = ch.get("target", {})
            if tgt.get("path") != "weights":
                continue
            node = self.gl["nodes"][tgt["node"]]
            if "mesh" not in node:
                continue
            s = anim["samplers"][ch["sampler"]]
            if s.get("interpolation", "LINEAR") != "LINEAR":
                continue
            times = self.accessor(s["input"]).astype(f32).reshape(-1)
            w = self.accessor(s["output"]).astype(f32).reshape(-1)
            k = len(w) // len(times)
            chans.append((node["mesh"], times, w[: k * len(times)].reshape(len(times), k)))
        return chans
