"""Write small GLB files for the reader's tests (test infrastructure: a minimal glTF 2.0 binary
writer and a PNG encoder that uses every row filter)."""
import json
import struct
import zlib

import numpy as np


def png_bytes(img, filters=(0, 1, 2, 3, 4), palette=None, trns=None):
    """img: (h, w, c) uint8 with c in 1..4, or (h, w) palette indices when `palette` is given.
    Row r uses filter filters[r % len(filters)]."""
    img = np.asarray(img, np.uint8)
    if palette is not None:
        ctype, chans = 3, 1
        img = img.reshape(img.shape[0], img.shape[1], 1)
    else:
        chans = img.shape[2]
        ctype = {1: 0, 2: 4, 3: 2, 4: 6}[chans]
    h, w = img.shape[:2]
    rows = img.reshape(h, w * chans).astype(np.int32)
    raw = bytearray()
    prior = np.zeros(w * chans, np.int32)
    for r in range(h):
        f = filters[r % len(filters)]
        cur = rows[r]
        left = np.concatenate([np.zeros(chans, np.int32), cur[:-chans]])
        upleft = np.concatenate([np.zeros(chans, np.int32), prior[:-chans]])
        if f == 0:
            out = cur
        elif f == 1:
            out = cur - left
        elif f == 2:
            out = cur - prior
        elif f == 3:
            out = cur - ((left + prior) >> 1)
        else:
            p = left + prior - upleft
            pa, pb, pc = np.abs(p - left), np.abs(p - prior), np.abs(p - upleft)
            pred = np.where((pa <= pb) & (pa <= pc), left, np.where(pb <= pc, prior, upleft))
            out = cur - pred
        raw.append(f)
        raw += bytes((out & 255).astype(np.uint8))
        prior = cur

    def chunk(kind, body):
        return struct.pack(">I", len(body)) + kind + body + struct.pack(">I", zlib.crc32(kind + body) & 0xFFFFFFFF)

    out = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, ctype, 0, 0, 0))
    if palette is not None:
        out += chunk(b"PLTE", bytes(np.asarray(palette, np.uint8).reshape(-1)))
        if trns is not None:
            out += chunk(b"tRNS", bytes(np.asarray(trns, np.uint8)))
    out += chunk(b"IDAT", zlib.compress(bytes(raw)))
    return out + chunk(b"IEND", b"")


class GlbWriter:
    def __init__(self):
        self.bin = bytearray()
        self.gl = {"asset": {"version": "2.0"}, "buffers": [{}], "bufferViews": [], "accessors": [],
                   "meshes": [], "materials": [], "nodes": [], "scenes": [{"nodes": []}], "scene": 0}

    def _view(self, data):
        while len(self.bin) % 4:
            self.bin.append(0)
        self.gl["bufferViews"].append({"buffer": 0, "byteOffset": len(self.bin), "byteLength": len(data)})
        self.bin += data
        return len(self.gl["bufferViews"]) - 1

    def accessor(self, arr, kind=None):
        arr = np.ascontiguousarray(arr)
        ctype = {np.dtype(np.float32): 5126, np.dtype(np.uint32): 5125, np.dtype(np.uint16): 5123}[arr.dtype]
        nc = 1 if arr.ndim == 1 else arr.shape[1]
        kind = kind or {1: "SCALAR", 2: "VEC2", 3: "VEC3", 4: "VEC4"}[nc]
        self.gl["accessors"].append({"bufferView": self._view(arr.tobytes()), "componentType": ctype,
                                     "count": int(arr.shape[0]), "type": kind})
        return len(self.gl["accessors"]) - 1

    def sparse_accessor(self, count, idx, vals):
        idx = np.ascontiguousarray(idx, np.uint16)
        vals = np.ascontiguousarray(vals, np.float32)
        self.gl["accessors"].append({
            "componentType": 5126, "count": count, "type": "VEC3",
            "sparse": {"count": int(len(idx)),
                       "indices": {"bufferView": self._view(idx.tobytes()), "componentType": 5123},
                       "values": {"bufferView": self._view(vals.tobytes())}}})
        return len(self.gl["accessors"]) - 1

    def image(self, png):
        self.gl.setdefault("images", []).append({"bufferView": self._view(png), "mimeType": "image/png",
                                                 "name": f"image{len(self.gl.get('images', []))}"})
        self.gl.setdefault("textures", []).append({"source": len(self.gl["images"]) - 1})
        return len(self.gl["textures"]) - 1

    def material(self, base=(1, 1, 1, 1), metallic=None, rough=None, base_tex=None, normal_tex=None):
        pbr = {"baseColorFactor": list(base)}
        if metallic is not None:
            pbr["metallicFactor"] = metallic
        if rough is not None:
            pbr["roughnessFactor"] = rough
        if base_tex is not None:
            pbr["baseColorTexture"] = {"index": base_tex}
        m = {"pbrMetallicRoughness": pbr}
        if normal_tex is not None:
            m["normalTexture"] = {"index": normal_tex}
        self.gl["materials"].append(m)
        return len(self.gl["materials"]) - 1

    def mesh(self, prims, weights=None):
        """prims: list of dicts with pos, idx (u16 / u32), material, optional nrm, uv, tan,
        targets (list of dicts with pos / nrm / tan arrays or ('sparse', idx, vals))."""
        out = []
        for p in prims:
            at = {"POSITION": self.accessor(np.asarray(p["pos"], np.float32))}
            for key, name in (("nrm", "NORMAL"), ("uv", "TEXCOORD_0"), ("tan", "TANGENT")):
                if p.get(key) is not None:
                    at[name] = self.accessor(np.asarray(p[key], np.float32))
            d = {"attributes": at, "indices": self.accessor(np.asarray(p["idx"]).reshape(-1)),
                 "material": p["material"]}
            if p.get("targets"):
                d["targets"] = []
                for t in p["targets"]:
                    td = {}
                    for key, name in (("pos", "POSITION"), ("nrm", "NORMAL"), ("tan", "TANGENT")):
                        v = t.get(key)
                        if v is None:
                            continue
                        if isinstance(v, tuple) and v[0] == "sparse":
                            td[name] = self.sparse_accessor(len(p["pos"]), v[1], v[2])
                        else:
                            td[name] = self.accessor(np.asarray(v, np.float32))
                    d["targets"].append(td)
            out.append(d)
        m = {"primitives": out}
        if weights is not None:
            m["weights"] = list(weights)
        self.gl["meshes"].append(m)
        return len(self.gl["meshes"]) - 1

    def node(self, root=True, **kw):
        self.gl["nodes"].append(kw)
        i = len(self.gl["nodes"]) - 1
        if root:
            self.gl["scenes"][0]["nodes"].append(i)
        return i

    def point_light(self, color, intensity):
        ext = self.gl.setdefault("extensions", {}).setdefault("KHR_lights_punctual", {"lights": []})
        ext["lights"].append({"type": "point", "color": list(color), "intensity": intensity})
        return len(ext["lights"]) - 1

    def weights_animation(self, node, times, weights, interpolation="LINEAR"):
        anim = {"samplers": [{"input": self.accessor(np.asarray(times, np.float32)),
                              "output": self.accessor(np.asarray(weights, np.float32).reshape(-1)),
                              "interpolation": interpolation}],
                "channels": [{"sampler": 0, "target": {"node": node, "path": "weights"}}]}
        self.gl.setdefault("animations", []).append(anim)

    def bytes(self):
        while len(self.bin) % 4:
            self.bin.append(0)
        self.gl["buffers"][0]["byteLength"] = len(self.bin)
        js = json.dumps(self.gl).encode()
        js += b" " * ((4 - len(js) % 4) % 4)
        body = struct.pack("<II", len(js), 0x4E4F534A) + js + struct.pack("<II", len(self.bin), 0x004E4942) + bytes(self.bin)
        return b"glTF" + struct.pack("<II", 2, 12 + len(body)) + body
