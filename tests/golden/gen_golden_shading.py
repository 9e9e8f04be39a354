"""Write the shading-attribute golden vectors from the reference's own compiled TUs.

    make -C oracle ref && python tests/golden/gen_golden_shading.py

oracle/_ref/libalpine_ref_render.so holds the reference's shapes/mesh.cpp, animation.cpp,
environment_map.cpp and materials/*.cpp compiled unmodified (oracle/Makefile) behind
oracle/ref_render_harness.cpp.  For each case the harness builds a Mesh::Data, runs
Animation::update -> Mesh::update (morph targets), then Mesh::getIntersectionAttributes and
Material::getBaseColor at random hits; and EnvironmentMap::sample for random directions.
The fixture (tests/golden/shading_small.npz) holds only inputs and outputs.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "shading_small.npz")


def grid_mesh(n, rng):
    """(n+1)^2 vertices on a wavy sheet, 2 n^2 triangles."""
    ys, xs = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
    v = np.stack([xs / n, ys / n, 0.1 * np.sin(3 * xs / n) * np.cos(2 * ys / n)], -1).reshape(-1, 3)
    v += rng.normal(0, 0.01, v.shape)
    t = []
    for j in range(n):
        for i in range(n):
            a = j * (n + 1) + i
            t += [(a, a + 1, a + n + 2), (a, a + n + 2, a + n + 1)]
    return v.astype(np.float32), np.array(t, np.uint32)


def ptr(a):
    return None if a is None else a.ctypes.data


def run_case(L, c):
    nv, nt = len(c["verts"]), len(c["prims"])
    nq = len(c["q_prim"])
    out_v = np.empty((nv, 3), np.float32)
    ns, ss, ts = (np.empty((nq, 3), np.float32) for _ in range(3))
    uv, bc = np.empty((nq, 2), np.float32), np.empty((nq, 3), np.float32)
    g = lambda k: c.get(k)
    nn = 0 if g("normals") is None else len(c["normals"])
    nuv = 0 if g("uvs") is None else len(c["uvs"])
    ntg = 0 if g("tpos") is None else c["tpos"].shape[0]
    nk = 0 if g("key_times") is None else len(c["key_times"])
    bct, nrt = g("bc_tex"), g("n_tex")
    L.ref_mesh_attributes(
        ptr(c["verts"]), nv, ptr(c["prims"]), nt, ptr(g("normals")), nn, ptr(g("normal_prims")),
        ptr(g("tangents")), ptr(g("uvs")), nuv, ptr(g("uv_prims")), ntg, ptr(g("tpos")),
        ptr(g("tnrm")), ptr(g("ttan")), nk, ptr(g("key_times")), ptr(g("key_weights")),
        float(c.get("time", 0.0)), int(c["mat_kind"]), ptr(c["bc"]), ptr(c["alpha"]), ptr(bct),
        0 if bct is None else bct.shape[1], 0 if bct is None else bct.shape[0], ptr(nrt),
        0 if nrt is None else nrt.shape[1], 0 if nrt is None else nrt.shape[0],
        ptr(c["q_prim"]), ptr(c["q_b"]), nq, ptr(out_v), ptr(ns), ptr(ss), ptr(ts), ptr(uv), ptr(bc))
    return dict(out_verts=out_v, ns=ns, ss=ss, ts=ts, uv=uv, base_color=bc)


def queries(rng, nt, n):
    b = rng.random((n, 2)).astype(np.float32)
    flip = b.sum(1) > 1
    b[flip] = 1 - b[flip]
    b[:8] = [[0, 0], [1, 0], [0, 1], [0.5, 0.5], [0.25, 0.25], [1e-7, 0], [0, 1 - 1e-7], [0.3, 0.1]]
    return rng.integers(0, nt, n).astype(np.uint32), b.astype(np.float32)


def main():
    L = oracle.ref_render_lib()
    rng = np.random.default_rng(77)
    f32 = lambda a: np.ascontiguousarray(a, np.float32)
    u32 = lambda a: np.ascontiguousarray(a, np.uint32)
    cases = {}
    # A: everything -- separate normal / uv index buffers, tangents, three morph targets under a
    # keyframed animation, base-colour and normal textures, uvs outside [0, 1]
    v, t = grid_mesh(8, rng)
    nv, nt = len(v), len(t)
    nn = nv + 7
    normal_prims = u32((t + 3) % nn)
    normals = f32(np.column_stack([rng.normal(0, 0.3, nn), rng.normal(0, 0.3, nn), 1 + rng.random(nn)]))
    tang = rng.normal(0, 1, (nn, 3))
    tang /= np.linalg.norm(tang, axis=1, keepdims=True)
    tangents = f32(np.column_stack([tang, rng.choice([-1.0, 1.0], nn)]))
    nuv = nv + 3
    uv_prims = u32((t * 7 + 1) % nuv)
    uvs = f32(rng.uniform(-1.5, 2.5, (nuv, 2)))
    ntg = 3
    q_prim, q_b = queries(rng, nt, 600)
    cases["a"] = dict(
        verts=v, prims=u32(t), normals=normals, normal_prims=normal_prims, tangents=tangents,
        uvs=uvs, uv_prims=uv_prims, tpos=f32(rng.normal(0, 0.05, (ntg, nv, 3))),
        tnrm=f32(rng.normal(0, 0.2, (ntg, nn, 3))), ttan=f32(rng.normal(0, 0.2, (ntg, nn, 3))),
        key_times=f32([0.0, 0.5, 1.3, 2.0]), key_weights=f32(rng.uniform(-0.3, 1.2, (4, ntg))),
        time=0.9, mat_kind=0, bc=f32([0.6, 0.5, 0.4]), alpha=f32([0.3, 0.3]),
        bc_tex=f32(rng.random((4, 8, 4))), n_tex=f32(rng.uniform(-1, 1, (4, 4, 4))),
        q_prim=q_prim, q_b=q_b)
    # B: the same animation clamped before the first key
    cases["b"] = dict(cases["a"], time=-1.0)
    # C: normals on the position indices and uvs with a base-colour texture, a conductor, no
    # tangents and no targets
    q_prim, q_b = queries(rng, nt, 400)
    cases["c"] = dict(verts=v, prims=u32(t), normals=f32(normals[:nv]), uvs=f32(uvs[:nv]),
                      mat_kind=1, bc=f32([0.9, 0.6, 0.5]), alpha=f32([0.2, 0.5]),
                      bc_tex=f32(rng.random((3, 5, 4))), q_prim=q_prim, q_b=q_b)
    # D: the morph weights themselves -- vertex j = 0 + 1 * w_j on x, so the blended vertices
    # read back as the weights of Animation::update (std::lerp of the bracketing keys)
    ntg = 5
    vd = np.zeros((ntg, 3), np.float32)
    tpos = np.zeros((ntg, ntg, 3), np.float32)
    for j in range(ntg):
        tpos[j, j, 0] = 1.0
    for k, tm in enumerate([0.0, 0.25, 0.7, 1.0, 1.7, 2.4]):
        cases[f"d{k}"] = dict(verts=vd, prims=u32([[0, 1, 2]]), tpos=f32(tpos),
                              key_times=f32([0.1, 0.6, 1.0, 2.0]),
                              key_weights=f32([[0, 1, -2, 0.5, 3e-8], [1, 1, 4, -0.5, 0.7],
                                               [0.25, -1, 4, 0.5, 1], [1, 0, 0, 0.5, -3]]),
                              time=tm, mat_kind=0, bc=f32([1, 1, 1]), alpha=f32([1, 1]),
                              q_prim=u32([0]), q_b=f32([[0.2, 0.3]]))
    out = {}
    for name, c in cases.items():
        for k, val in c.items():
            if val is not None:
                out[f"{name}_{k}"] = np.asarray(val)
        for k, val in run_case(L, c).items():
            out[f"{name}_out_{k}"] = val
    # environment map
    w, h = 16, 8
    env = f32(rng.random((h, w, 3)))
    d = rng.normal(0, 1, (1000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d[:10] = [[0, 1, 0], [0, -1, 0], [-1, 0, 0], [1, 0, 0], [0, 0, 1], [0, 0, -1],
              [-1, 0, 1e-9], [-1, 0, -1e-9], [0.6, 0.8, 0], [0, 1.0000001, 0]]
    d = f32(d)
    e = np.empty_like(d)
    L.ref_environment_sample(env.ctypes.data, w, h, d.ctypes.data, len(d), e.ctypes.data)
    out.update(env_map=env, env_dirs=d, env_out=e)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, sum(v.nbytes for v in out.values()), "bytes")


if __name__ == "__main__":
    main()
