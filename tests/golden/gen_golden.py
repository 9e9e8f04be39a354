"""Write the traversal golden vectors from the reference's own compiled BVH.

Run in the container that holds the reference tree:

    make -C oracle ref && python tests/golden/gen_golden.py

The reference's bvh.cpp / bvh_common.cpp (compiled by oracle/Makefile into
oracle/_ref/libalpine_ref.so) build the binary BVH over a small procedural
scene and trace fixed rays; the scene, the rays and the results are stored as
data in tests/golden/traversal_small.npz.  Nothing from the reference's source
is copied; the fixture holds inputs and outputs only.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from alpine_amd import scenes  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "traversal_small.npz")


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/libalpine_ref.so is missing: run `make -C oracle ref`")
    verts, tris = scenes.small_scene()
    lo, hi = scenes.scene_bounds(verts)
    o_inc, d_inc = scenes.incoherent_rays(3000, lo, hi, seed=8)
    o_coh, d_coh = scenes.coherent_rays(32, 32, eye=(0.0, 0.2, -3.5), fovy_deg=60.0)
    # axis-aligned and near-axis directions (zero components in the slab test)
    rng = np.random.default_rng(11)
    o_ax = rng.uniform(lo, hi, (600, 3)).astype(np.float32)
    axes = np.eye(3, dtype=np.float32)
    d_ax = np.concatenate([axes, -axes])[rng.integers(0, 6, 600)]
    org = np.concatenate([o_inc, o_coh, o_ax]).astype(np.float32)
    dir = np.concatenate([d_inc, d_coh, d_ax]).astype(np.float32)
    tfar = rng.uniform(0.0, float(np.linalg.norm(hi - lo)), len(org)).astype(np.float32)

    ref = oracle.RefBvh(verts, tris)
    h = ref.intersect(org, dir)
    occ = ref.occluded(org, dir)
    occ_t = ref.occluded(org, dir, tfar)
    st = oracle.ref_tree_stats(verts, tris, 8)
    st16 = oracle.ref_tree_stats(verts, tris, 16)
    np.savez_compressed(
        OUT, verts=verts, tris=tris, org=org, dir=dir, tfar=tfar,
        prim=h.prim, t=h.t, b0=h.b0, b1=h.b1, occ=occ, occ_tfar=occ_t,
        leaf_keys8=st["leaf_keys"], leaf_keys16=st16["leaf_keys"],
        sah8=np.float64(st["sah_cost"]), sah_leaf8=np.float64(st["sah_cost_leaf"]),
        nodes8=np.uint32(st["node_count"]), leaves8=np.uint32(st["leaf_count"]),
        sah16=np.float64(st16["sah_cost"]), nodes16=np.uint32(st16["node_count"]),
    )
    print("wrote", OUT, os.path.getsize(OUT), "bytes;", int((h.prim != 0xFFFFFFFF).sum()),
          "hits of", len(org))


if __name__ == "__main__":
    main()
