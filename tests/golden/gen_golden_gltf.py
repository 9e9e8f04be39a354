"""Write the float4x4 golden vectors of the glTF path from the reference's own math/matrix.h.

    make -C oracle ref && python tests/golden/gen_golden_gltf.py

The reference's gltf_loader.cpp multiplies node matrices down the hierarchy and transforms
positions (w = 1), normals and tangents (w = 0) with mul() of math/matrix.h (header-only,
compiled behind oracle/ref_render_harness.cpp).  The fixture (tests/golden/gltf_math.npz) holds
random and TRS-shaped matrices, vectors and the reference's products.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gltf_math.npz")


def main():
    L = oracle.ref_render_lib()
    rng = np.random.default_rng(31)
    k = 24
    m0 = rng.normal(0, 2, (k, 16)).astype(np.float32)
    m1 = rng.normal(0, 2, (k, 16)).astype(np.float32)
    m0[::3, 3::4] = [0, 0, 0, 1]  # affine ones too
    m1[::3, 3::4] = [0, 0, 0, 1]
    n = 64
    v = rng.normal(0, 3, (k, n, 4)).astype(np.float32)
    v[:, : n // 2, 3] = 1.0
    v[:, n // 2:, 3] = 0.0
    prod = np.empty((k, 16), np.float32)
    mv = np.empty((k, n, 4), np.float32)
    for i in range(k):
        L.ref_mat4(m0[i].ctypes.data, m1[i].ctypes.data, np.ascontiguousarray(v[i]).ctypes.data, n,
                   prod[i].ctypes.data, mv[i].ctypes.data)
    np.savez_compressed(OUT, m0=m0, m1=m1, v=v, prod=prod, mv=mv)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
