"""Write the render-path golden vectors from the reference's own compiled TUs.

    make -C oracle ref && python tests/golden/gen_golden_render.py

oracle/_ref/libalpine_ref_render.so is the reference's sampler.cpp,
camera.cpp, utils/util.cpp, utils/alias_table.cpp, materials/*.cpp,
lights/{disk,point}_light.cpp and light_samplers/*.cpp compiled unmodified
(oracle/Makefile) behind oracle/ref_render_harness.cpp.  The fixture
(tests/golden/render_small.npz) holds only inputs and outputs.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "render_small.npz")

MATERIALS = [  # kind, base colour, alpha
    (0, (0.73, 0.73, 0.73), (1.0, 1.0)),      # Lambert-only Dielectric (OBJ path)
    (0, (0.65, 0.05, 0.05), (0.3, 0.3)),      # Dielectric with GGX coat
    (0, (0.2, 0.6, 0.3), (0.1, 0.45)),        # anisotropic coat
    (1, (0.95, 0.64, 0.54), (0.05, 0.05)),    # rough conductors
    (1, (0.9, 0.9, 0.9), (0.4, 0.8)),
]


def unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def main():
    L = oracle.ref_render_lib()
    rng = np.random.default_rng(2024)
    out = {}
    # sampler sequences
    seeds = np.array([0, 1, 7, 65535, 123456789], np.uint32)
    seq = np.empty((len(seeds), 64), np.float32)
    for i, s in enumerate(seeds):
        L.ref_sampler_sequence(int(s), 64, seq[i])
    out["sampler_seeds"], out["sampler_seq"] = seeds, seq
    # camera
    xy = rng.random((256, 2)).astype(np.float32)
    eye, tgt, up = (np.float32([0.3, 0.5, -3.0]), np.float32([0.0, 0.2, 0.0]),
                    np.float32([0.0, 1.0, 0.0]))
    org, dir = np.empty((256, 3), np.float32), np.empty((256, 3), np.float32)
    L.ref_camera_rays(eye, tgt, up, np.float32(0.7), np.float32(1.25), xy, 256, org, dir)
    out.update(cam_eye=eye, cam_target=tgt, cam_up=up, cam_fovy=np.float32(0.7),
               cam_aspect=np.float32(1.25), cam_xy=xy, cam_org=org, cam_dir=dir)
    # materials
    n = 2000
    wo = unit(rng.standard_normal((n, 3))).astype(np.float32)
    wo[: n // 10 * 9, 2] = np.abs(wo[: n // 10 * 9, 2])  # mostly upper hemisphere
    u = rng.random((n, 2)).astype(np.float32)
    wi_eval = unit(rng.standard_normal((n, 3))).astype(np.float32)
    for k, (kind, bc, alpha) in enumerate(MATERIALS):
        bc = np.float32(bc)
        al = np.float32(alpha)
        ok = np.empty(n, np.uint8)
        est, wi = np.empty((n, 3), np.float32), np.empty((n, 3), np.float32)
        pdf = np.empty(n, np.float32)
        L.ref_material_sample(kind, bc, al, wo, u, n, ok, est, wi, pdf)
        f, fp = np.empty((n, 3), np.float32), np.empty(n, np.float32)
        L.ref_material_eval(kind, bc, al, wo, wi_eval, n, f, fp)
        out.update({f"mat{k}_ok": ok, f"mat{k}_est": est, f"mat{k}_wi": wi, f"mat{k}_pdf": pdf,
                    f"mat{k}_f": f, f"mat{k}_fpdf": fp})
    out["mat_specs"] = np.array([[kind, *bc, *alpha] for kind, bc, alpha in MATERIALS], np.float32)
    out["mat_wo"], out["mat_u"], out["mat_wi_eval"] = wo, u, wi_eval
    # lights: 37 disk + point lights, log-normal power
    nl = 37
    spec = np.zeros((nl, 12), np.float32)
    spec[:, 0] = (rng.random(nl) < 0.2).astype(np.float32)  # 1 = point
    spec[:, 1:4] = rng.uniform(-4, 4, (nl, 3))
    spec[:, 4:7] = unit(rng.standard_normal((nl, 3)))
    spec[:, 7:10] = rng.uniform(0.2, 1.0, (nl, 3))
    spec[:, 10] = np.exp(rng.normal(0.0, 1.5, nl))
    spec[:, 11] = rng.uniform(0.05, 0.5, nl)
    m = 3000
    hit = rng.uniform(-5, 5, (m, 3)).astype(np.float32)
    ns = unit(rng.standard_normal((m, 3))).astype(np.float32)
    uu = rng.random(m).astype(np.float32)
    for kind in (0, 1, 2):
        idx = np.empty(m, np.int32)
        pdf = np.empty(m, np.float32)
        L.ref_light_sampler(kind, spec, nl, uu, hit, ns, m, idx, pdf)
        out[f"ls{kind}_idx"], out[f"ls{kind}_pdf"] = idx, pdf
    out["ls_spec"], out["ls_hit"], out["ls_ns"], out["ls_u"] = spec, hit, ns, uu
    # one disk light: sample + pdf
    disk = spec[~spec[:, 0].astype(bool)][0:1].copy()
    u2 = rng.random((m, 2)).astype(np.float32)
    ok = np.empty(m, np.uint8)
    le, wi = np.empty((m, 3), np.float32), np.empty((m, 3), np.float32)
    dist, pdf = np.empty(m, np.float32), np.empty(m, np.float32)
    L.ref_light_sample(disk, u2, hit, m, ok, le, wi, dist, pdf)
    wq = unit(rng.standard_normal((m, 3))).astype(np.float32)
    # aim half the queries at the disk so computePdf has hits
    tgt_pts = disk[0, 1:4] + 0.5 * disk[0, 11] * unit(rng.standard_normal((m, 3)))
    wq[: m // 2] = unit(tgt_pts[: m // 2] - hit[: m // 2])
    qpdf, qdist = np.empty(m, np.float32), np.empty(m, np.float32)
    L.ref_light_pdf(disk, hit, wq, m, qpdf, qdist)
    out.update(disk_spec=disk, disk_u=u2, disk_ok=ok, disk_le=le, disk_wi=wi, disk_dist=dist,
               disk_pdf=pdf, disk_q=wq, disk_qpdf=qpdf, disk_qdist=qdist)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
