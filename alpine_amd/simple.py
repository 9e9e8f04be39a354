"""`simple`, the reference's command-line renderer (applications/simple/main.cpp), on the GPU.

    python -m alpine_amd.simple -i scene.obj -o image.ppm --spp 64 \\
        --accelerator gpu --lightSampler bvh

Same fixed setup as the reference app (main.cpp:104-128): 256 x 256, max depth 8, a white
background, one disk light of power 100 at (0, 10, 0) facing the origin with radius 5, a camera
at (0, 0, -3) looking at the origin with a 90 degree field of view, and postProcess with the
ACES tone map (main.cpp:157).  Scenes are read by this module's own OBJ/MTL reader (the
reference uses tinyobjloader, obj_loader.cpp:16-122): every face's material is the MTL `Kd`
as a Lambert-only Dielectric (alpha 1, obj_loader.cpp:76-78); polygons are fanned into
triangles; all objects form one mesh.  `.glb` files go through alpine_amd.gltf (the
reference's glTF path): meshes with normals, tangents, uvs and morph targets, PBR materials
with textures, punctual lights, and morph animations applied at time 0 before the build
(main.cpp:137-140).  Output: `.ppm` (P3, like image.cpp:7-33) or `.pfm` (the linear mean
radiance, little-endian floats).
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

WIDTH = HEIGHT = 256
MAX_DEPTH = 8


def _index(tok, n):
    """OBJ 1-based (or negative, relative) index -> 0-based."""
    i = int(tok)
    return i - 1 if i > 0 else n + i


def read_mtl(path):
    """newmtl / Kd of an MTL file -> {name: (r, g, b)}."""
    mats, cur = {}, None
    with open(path, "r", encoding="utf-8", errors="replace") as f:
        for line in f:
            parts = line.split()
            if not parts or parts[0].startswith("#"):
                continue
            if parts[0] == "newmtl":
                cur = " ".join(parts[1:])
                mats[cur] = (0.8, 0.8, 0.8)
            elif parts[0] == "Kd" and cur is not None and len(parts) >= 4:
                mats[cur] = tuple(float(x) for x in parts[1:4])
    return mats


def read_obj(path):
    """Vertices, triangles, per-triangle material index and the material colours of an OBJ."""
    verts, tris, tri_mat = [], [], []
    names, colours = {}, []
    library = {}
    current = None
    base = os.path.dirname(os.path.abspath(path))

    def material_id(name):
        if name not in names:
            names[name] = len(colours)
            colours.append(library.get(name, (0.8, 0.8, 0.8)))
        return names[name]

    with open(path, "r", encoding="utf-8", errors="replace") as f:
        for line in f:
            parts = line.split()
            if not parts or parts[0].startswith("#"):
                continue
            key = parts[0]
            if key == "v":
                verts.append([float(x) for x in parts[1:4]])
            elif key == "f":
                idx = [_index(p.split("/")[0], len(verts)) for p in parts[1:]]
                mid = material_id(current)
                for k in range(1, len(idx) - 1):
                    tris.append((idx[0], idx[k], idx[k + 1]))
                    tri_mat.append(mid)
            elif key == "mtllib":
                for lib in parts[1:]:
                    p = os.path.join(base, lib)
                    if os.path.exists(p):
                        library.update(read_mtl(p))
            elif key == "usemtl":
                current = " ".join(parts[1:])
    if not colours:
        colours.append((0.8, 0.8, 0.8))
    return (np.asarray(verts, np.float32).reshape(-1, 3), np.asarray(tris, np.uint32).reshape(-1, 3),
            np.asarray(tri_mat, np.uint32), colours)


def write_ppm(path, frame):
    """P3 text image as image.cpp:7-33 writes it (frame: h x w x 3 uint8)."""
    h, w, _ = frame.shape
    with open(path, "w") as f:
        f.write(f"P3\n{w} {h}\n255\n")
        for row in frame.reshape(h, w * 3):
            for x in range(w):
                f.write(f"{row[3 * x]} {row[3 * x + 1]} {row[3 * x + 2]}\n")


def write_pfm(path, rgb):
    """Linear RGB as PFM (top row last, little-endian)."""
    h, w, _ = rgb.shape
    with open(path, "wb") as f:
        f.write(f"PF\n{w} {h}\n-1.0\n".encode())
        f.write(np.ascontiguousarray(rgb[::-1], "<f4").tobytes())


def load_glb_scene(acc, sc, time_s=0.0):
    """Put a GLB scene (alpine_amd.gltf.read_glb) on the accelerator as alpine's loadScene +
    updateScene(time) do: meshes in mesh order with their shading attributes and morph targets,
    the file's textures, animations applied at `time_s` (the caller builds afterwards).  Returns
    (per-triangle material indices, materials, point lights)."""
    from alpine_amd.animation import MorphAnimation
    shape_of, tri_mat = {}, []
    for mi in sorted(sc.meshes):
        me = sc.meshes[mi]
        sid = acc.append_mesh(me["verts"], me["tris"])
        shape_of[mi] = sid
        acc.set_shading(sid, me["normals"], None, me["tangents"], me["uvs"])
        tg = me["targets"]
        if tg:
            nv = len(me["verts"])
            z = np.zeros((nv, 3), np.float32)
            dpos = np.stack([t["pos"] if t["pos"] is not None else z for t in tg])
            dn = dt = None
            if me["normals"] is not None:
                dn = np.stack([t["nrm"] if t["nrm"] is not None else z for t in tg])
                dt = np.stack([t["tan"] if t["tan"] is not None else z for t in tg])
            acc.set_targets(sid, dpos, dn, dt)
        tri_mat.append(me["tri_mat"])
    for anim in sc.animations:
        MorphAnimation([(shape_of[mi], t, w) for mi, t, w in anim if mi in shape_of]).update(acc, time_s)
    acc.set_textures(sc.textures)
    tm = np.concatenate(tri_mat) if tri_mat else np.zeros(0, np.uint32)
    return tm, sc.materials, sc.lights


def main(argv=None):
    ap = argparse.ArgumentParser(prog="simple", description=__doc__.split("\n")[0])
    ap.add_argument("-i", "--input", required=True, help="input file (obj or glb)")
    ap.add_argument("-o", "--output", required=True, help="output image (ppm or pfm)")
    ap.add_argument("--spp", type=int, required=True)
    ap.add_argument("--accelerator", required=True, choices=["gpu"])
    ap.add_argument("--lightSampler", required=True, choices=["uniform", "power", "bvh"])
    ap.add_argument("--denoiser", action="store_true", help="not available (no OIDN)")
    ap.add_argument("--environment", help="Radiance .hdr environment map seen on a miss "
                    "(Alpine::loadEnvironmentMap, alpine.cpp:253-257); default: white background")
    a = ap.parse_args(argv)
    ext = os.path.splitext(a.input)[1]
    if ext not in (".obj", ".glb"):
        print("ERROR: File extension must be .obj or .glb.")
        return 1
    if os.path.splitext(a.output)[1] not in (".ppm", ".pfm"):
        print("ERROR: output must be .ppm or .pfm")
        return 1
    import alpine_amd as aa
    from alpine_amd import gltf
    point_lights = []
    try:
        if ext == ".obj":
            scene = read_obj(a.input)
        else:
            scene = gltf.read_glb(a.input)
            for m in scene.messages:
                print(m)
    except (OSError, ValueError, KeyError, IndexError, gltf.GltfError) as e:
        print(f"ERROR: failed to load {a.input}: {e}")
        return 1
    env = None
    if a.environment:
        from alpine_amd import hdr
        try:
            env = hdr.read_hdr(a.environment)
        except (OSError, ValueError) as e:
            print(f"ERROR: failed to load {a.environment}: {e}")
            return 1
    acc = aa.GpuAccelerator(0)
    if env is not None:
        acc.set_environment(env)
    if ext == ".obj":
        verts, tris, tri_mat, colours = scene
        acc.append_mesh(verts, tris)
        mats = [aa.material(aa.MAT_DIELECTRIC, c, (1.0, 1.0)) for c in colours]
    else:
        tri_mat, gmats, point_lights = load_glb_scene(acc, scene, 0.0)
        mats = [aa.material(k, bc, al, bt, nt) for k, bc, al, bt, nt in gmats]
    t0 = time.perf_counter()
    acc.update_scene()
    print(f"Accelerator Build: {(time.perf_counter() - t0) * 1e3:.3f} [ms]")
    acc.set_materials(mats, tri_mat)
    sampler = {"uniform": aa.LS_UNIFORM, "power": aa.LS_POWER, "bvh": aa.LS_BVH}[a.lightSampler]
    pos, target = np.array([0.0, 10.0, 0.0]), np.zeros(3)
    normal = (target - pos) / np.linalg.norm(target - pos)
    t0 = time.perf_counter()
    # the app's disk light first, then the file's point lights (main.cpp:125, gltf_loader.cpp:460-472)
    lights = [aa.light(aa.LIGHT_DISK, p0=pos, normal=normal, color=(1.0, 1.0, 1.0), power=100.0, radius=5.0)]
    lights += [aa.light(aa.LIGHT_POINT_INTENSITY, p0=p, color=i) for i, p in point_lights]
    acc.set_lights(lights, sampler)
    print(f"Light Sampler Build: {(time.perf_counter() - t0) * 1e3:.3f} [ms]")
    acc.set_camera(aa.camera((0.0, 0.0, -3.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), np.pi / 2,
                             WIDTH / HEIGHT))
    acc.render_reset(WIDTH, HEIGHT, 0)
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    acc.render(a.spp, MAX_DEPTH, 0, (1.0, 1.0, 1.0))
    torch.cuda.synchronize()
    print(f"Render: {(time.perf_counter() - t0) * 1e3:.3f} [ms]")
    if a.output.endswith(".pfm"):
        rgb, _, n = acc.render_read()
        write_pfm(a.output, rgb / n)
    else:
        write_ppm(a.output, acc.postprocess(tonemap=True).cpu().numpy())
    return 0


if __name__ == "__main__":
    sys.exit(main())
