"""Procedural scenes and ray sets for the traversal microbench and the parity tests.

Every generator is deterministic (numpy PCG64 with a fixed seed), so the same
arrays are produced on any host.  Geometry is returned as (verts float32
(nv, 3), tris uint32 (nt, 3)); ray sets as (org float32 (n, 3), dir float32
(n, 3)).

Workloads
---------
* ``displaced_icosphere(subdiv=8, seed=7)``: icosphere (20 * 4**subdiv
  triangles) whose vertices are pushed along the normal to radius
  1 + 0.1 * value_noise(p), value noise in [-1, 1] on an integer lattice of
  frequency 4 (smoothstep-blended trilinear interpolation of hashed lattice
  values).
* ``random_triangles(n=200_000, seed=7)``: equilateral triangles, centres
  uniform in [-3, 3]^3, edge length log-uniform in [0.005, 0.1], uniformly
  random orientation.
* ``microbench_scene()``: the two above concatenated (1,510,720 triangles at
  the default sizes).
* ``incoherent_rays(n, lo, hi, seed)``: origins uniform in the box, directions
  uniform on the sphere (z = 2u - 1, phi = 2 pi v).
* ``coherent_rays(w, h, ...)``: pinhole camera rays through pixel centres.
"""
from __future__ import annotations

import numpy as np

_PHI = (1.0 + 5.0 ** 0.5) / 2.0


def icosahedron():
    v = np.array([
        [-1, _PHI, 0], [1, _PHI, 0], [-1, -_PHI, 0], [1, -_PHI, 0],
        [0, -1, _PHI], [0, 1, _PHI], [0, -1, -_PHI], [0, 1, -_PHI],
        [_PHI, 0, -1], [_PHI, 0, 1], [-_PHI, 0, -1], [-_PHI, 0, 1],
    ], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([
        [0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
        [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
        [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
        [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1],
    ], dtype=np.int64)
    return v, f


def subdivide(v, f):
    """Split every triangle into four; edge midpoints are projected to the unit sphere."""
    nv = len(v)
    e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]], axis=0)
    e.sort(axis=1)
    key = e[:, 0] * nv + e[:, 1]
    uniq, inv = np.unique(key, return_inverse=True)
    a, b = uniq // nv, uniq % nv
    mid = v[a] + v[b]
    mid /= np.linalg.norm(mid, axis=1, keepdims=True)
    m = nv + inv.reshape(3, -1).T  # (nf, 3): midpoints of edges 01, 12, 20
    v2 = np.concatenate([v, mid], axis=0)
    f0, f1, f2 = f[:, 0], f[:, 1], f[:, 2]
    m01, m12, m20 = m[:, 0], m[:, 1], m[:, 2]
    f2_ = np.concatenate([
        np.stack([f0, m01, m20], 1), np.stack([m01, f1, m12], 1),
        np.stack([m20, m12, f2], 1), np.stack([m01, m12, m20], 1),
    ], axis=0)
    return v2, f2_


def _hash3(ix, iy, iz, seed):
    h = (ix.astype(np.uint64) * np.uint64(0x9E3779B185EBCA87)
         ^ iy.astype(np.uint64) * np.uint64(0xC2B2AE3D27D4EB4F)
         ^ iz.astype(np.uint64) * np.uint64(0x165667B19E3779F9)
         ^ np.uint64(seed) * np.uint64(0x27D4EB2F165667C5))
    h ^= h >> np.uint64(33)
    h *= np.uint64(0xFF51AFD7ED558CCD)
    h ^= h >> np.uint64(33)
    h *= np.uint64(0xC4CEB9FE1A85EC53)
    h ^= h >> np.uint64(33)
    return (h >> np.uint64(11)).astype(np.float64) / float(1 << 53)


def value_noise(p, seed=7, freq=4.0):
    """Trilinear value noise in [-1, 1] with smoothstep weights."""
    q = p * freq
    i0 = np.floor(q).astype(np.int64)
    fr = q - i0
    w = fr * fr * (3.0 - 2.0 * fr)
    out = np.zeros(len(p))
    with np.errstate(over="ignore"):
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    val = _hash3(i0[:, 0] + dx, i0[:, 1] + dy, i0[:, 2] + dz, seed)
                    wx = w[:, 0] if dx else 1.0 - w[:, 0]
                    wy = w[:, 1] if dy else 1.0 - w[:, 1]
                    wz = w[:, 2] if dz else 1.0 - w[:, 2]
                    out += wx * wy * wz * val
    return out * 2.0 - 1.0


def displaced_icosphere(subdiv=8, seed=7, amp=0.1):
    v, f = icosahedron()
    for _ in range(subdiv):
        v, f = subdivide(v, f)
    r = 1.0 + amp * value_noise(v, seed)
    v = v * r[:, None]
    return v.astype(np.float32), f.astype(np.uint32)


def _random_rotations(rng, n):
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1),
    ], axis=1)


def random_triangles(n=200_000, seed=7, extent=3.0, lmin=0.005, lmax=0.1):
    rng = np.random.default_rng(seed + 1000)
    c = rng.uniform(-extent, extent, (n, 3))
    L = np.exp(rng.uniform(np.log(lmin), np.log(lmax), n))
    R = _random_rotations(rng, n)
    ang = np.array([0.0, 2.0 * np.pi / 3.0, 4.0 * np.pi / 3.0])
    local = np.stack([np.cos(ang), np.sin(ang), np.zeros(3)], -1) / np.sqrt(3.0)  # unit side
    pts = c[:, None, :] + L[:, None, None] * np.einsum("nij,kj->nki", R, local)
    verts = pts.reshape(-1, 3).astype(np.float32)
    tris = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    return verts, tris


def concat(*meshes):
    vs, ts, off = [], [], 0
    for v, t in meshes:
        vs.append(v)
        ts.append(t + np.uint32(off))
        off += len(v)
    return np.concatenate(vs).astype(np.float32), np.concatenate(ts).astype(np.uint32)


def microbench_scene(subdiv=8, n_random=200_000, seed=7):
    return concat(displaced_icosphere(subdiv, seed), random_triangles(n_random, seed))


def scene_bounds(verts):
    return verts.min(axis=0), verts.max(axis=0)


def incoherent_rays(n, lo, hi, seed=8):
    rng = np.random.default_rng(seed)
    org = rng.uniform(np.asarray(lo, np.float64), np.asarray(hi, np.float64), (n, 3))
    u, v = rng.random(n), rng.random(n)
    z = 2.0 * u - 1.0
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    phi = 2.0 * np.pi * v
    d = np.stack([r * np.cos(phi), r * np.sin(phi), z], -1)
    return org.astype(np.float32), d.astype(np.float32)


def coherent_rays(width=2048, height=2048, eye=(0.0, 0.0, -6.5), target=(0.0, 0.0, 0.0),
                  up=(0.0, 1.0, 0.0), fovy_deg=50.0):
    eye = np.asarray(eye, np.float64)
    fwd = np.asarray(target, np.float64) - eye
    fwd /= np.linalg.norm(fwd)
    right = np.cross(np.asarray(up, np.float64), fwd)
    right /= np.linalg.norm(right)
    upv = np.cross(fwd, right)
    h = np.tan(np.radians(fovy_deg) * 0.5)
    w = h * width / height
    xs = ((np.arange(width) + 0.5) / width - 0.5) * 2.0 * w
    ys = (0.5 - (np.arange(height) + 0.5) / height) * 2.0 * h
    X, Y = np.meshgrid(xs, ys)
    d = X[..., None] * right + Y[..., None] * upv + fwd
    d = d.reshape(-1, 3)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    org = np.broadcast_to(eye, d.shape)
    return np.ascontiguousarray(org, np.float32), d.astype(np.float32)


def small_scene(n_random=2000, subdiv=3, seed=7):
    """A few thousand triangles with the microbench's structure, for brute-force checks."""
    return concat(displaced_icosphere(subdiv, seed), random_triangles(n_random, seed,
                                                                      extent=1.5, lmin=0.02,
                                                                      lmax=0.4))


def geometric_chain(n=852, ratio=1.05, size=1e-6):
    """A deliberately deep tree: n tiny right triangles at x = ratio^i along the x axis
    (triangle i spans x_i .. x_i (1 + size) and 0 .. size x_i in y and z).  Binned SAH peels a
    few triangles off the far end per level, so the binary tree is about 30 levels deep and
    the 8-wide tree 24 (at the defaults: x up to 2^60, where surface areas stay finite)."""
    x = ratio ** np.arange(n, dtype=np.float64)
    v = np.zeros((3 * n, 3), np.float64)
    v[0::3, 0] = x
    v[1::3, 0] = x * (1.0 + size)
    v[1::3, 1] = size * x
    v[2::3, 0] = x
    v[2::3, 2] = size * x
    return v.astype(np.float32), np.arange(3 * n, dtype=np.uint32).reshape(-1, 3)


def chain_rays(n, verts, seed=9):
    """Rays through the chain's boxes: origins on the chain's axis below its first triangle,
    directions within 0.01 rad of +x, offsets in y and z up to the first triangles' size."""
    rng = np.random.default_rng(seed)
    x0 = float(verts[:, 0].min())
    s = float(verts[1, 1])
    org = np.stack([np.full(n, 0.5 * x0), rng.uniform(0.0, s, n), rng.uniform(0.0, s, n)], -1)
    a = rng.uniform(-0.01, 0.01, (n, 2))
    d = np.stack([np.ones(n), a[:, 0], a[:, 1]], -1)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return org.astype(np.float32), d.astype(np.float32)


# ---------------------------------------------------------------------------
# Render scenes
# ---------------------------------------------------------------------------
def quad_mesh(origin, u, v, nu, nv):
    """Tessellated parallelogram; triangle normals point along cross(u, v)."""
    origin, u, v = (np.asarray(a, np.float64) for a in (origin, u, v))
    a = np.linspace(0.0, 1.0, nu + 1)
    b = np.linspace(0.0, 1.0, nv + 1)
    A, B = np.meshgrid(a, b, indexing="ij")
    verts = origin + A[..., None] * u + B[..., None] * v
    idx = np.arange((nu + 1) * (nv + 1)).reshape(nu + 1, nv + 1)
    v00, v10 = idx[:-1, :-1].ravel(), idx[1:, :-1].ravel()
    v11, v01 = idx[1:, 1:].ravel(), idx[:-1, 1:].ravel()
    tris = np.concatenate([np.stack([v00, v10, v11], 1), np.stack([v00, v11, v01], 1)])
    return verts.reshape(-1, 3).astype(np.float32), tris.astype(np.uint32)


def box_mesh(sx, h, sz, angle_deg, pos, n):
    """Closed box [-sx/2, sx/2] x [0, h] x [-sz/2, sz/2], outward normals, rotated about y."""
    faces = [
        ((sx / 2, 0, -sz / 2), (0, h, 0), (0, 0, sz)),
        ((-sx / 2, 0, -sz / 2), (0, 0, sz), (0, h, 0)),
        ((-sx / 2, h, -sz / 2), (0, 0, sz), (sx, 0, 0)),
        ((-sx / 2, 0, -sz / 2), (sx, 0, 0), (0, 0, sz)),
        ((-sx / 2, 0, sz / 2), (sx, 0, 0), (0, h, 0)),
        ((-sx / 2, 0, -sz / 2), (0, h, 0), (sx, 0, 0)),
    ]
    mesh = concat(*[quad_mesh(o, u, v, n, n) for o, u, v in faces])
    t = np.radians(angle_deg)
    R = np.array([[np.cos(t), 0, np.sin(t)], [0, 1, 0], [-np.sin(t), 0, np.cos(t)]])
    verts = mesh[0].astype(np.float64) @ R.T + np.asarray(pos, np.float64)
    return verts.astype(np.float32), mesh[1]


class RenderScene:
    """Geometry + per-triangle material + lights + camera for the path tracer."""

    def __init__(self, verts, tris, tri_mat, materials, lights, camera, width, height, spp,
                 max_depth, light_sampler, background=(0.0, 0.0, 0.0), rr_depth=0, seed=0):
        self.verts, self.tris, self.tri_mat = verts, tris, np.asarray(tri_mat, np.uint32)
        self.materials = materials    # list of (kind, base_color, alpha)
        self.lights = lights          # list of dicts for alpine_amd.light(**d)
        self.camera = camera          # dict: eye, target, up, fovy (radians), aspect
        self.width, self.height, self.spp, self.max_depth = width, height, spp, max_depth
        self.light_sampler = light_sampler
        self.background = background
        self.rr_depth = rr_depth
        self.seed = seed


def cornell_scene(width=256, height=256, spp=16, max_depth=5, seed=1, wall_res=32, box_res=8):
    """Cornell-style room: 5 walls (32 x 32 quads each) + two boxes rotated -17 / +17 degrees
    (8 x 8 per face) -> 11,776 triangles; one 0.25 x 0.25 quad light of radiance (17, 12, 4)
    under the ceiling; Lambertian grey 0.73 / red (0.65, 0.05, 0.05) / green (0.12, 0.45, 0.15);
    pinhole camera, 40 degree vertical field of view; uniform light sampler."""
    walls = [
        (quad_mesh((-1, 0, -1), (0, 0, 2), (2, 0, 0), wall_res, wall_res), 0),  # floor
        (quad_mesh((-1, 2, -1), (2, 0, 0), (0, 0, 2), wall_res, wall_res), 0),  # ceiling
        (quad_mesh((-1, 0, 1), (0, 2, 0), (2, 0, 0), wall_res, wall_res), 0),   # back
        (quad_mesh((-1, 0, -1), (0, 2, 0), (0, 0, 2), wall_res, wall_res), 1),  # left, red
        (quad_mesh((1, 0, -1), (0, 0, 2), (0, 2, 0), wall_res, wall_res), 2),   # right, green
        (box_mesh(0.6, 0.6, 0.6, -17.0, (0.35, 0.0, -0.3), box_res), 0),
        (box_mesh(0.6, 1.2, 0.6, 17.0, (-0.35, 0.0, 0.35), box_res), 0),
    ]
    verts, tris = concat(*[m for m, _ in walls])
    tri_mat = np.concatenate([np.full(len(m[1]), k, np.uint32) for m, k in walls])
    materials = [(0, (0.73, 0.73, 0.73), (1.0, 1.0)), (0, (0.65, 0.05, 0.05), (1.0, 1.0)),
                 (0, (0.12, 0.45, 0.15), (1.0, 1.0))]
    lights = [dict(kind=2, p0=(-0.125, 1.999, -0.125), p1=(0.125, 1.999, -0.125),
                   p2=(-0.125, 1.999, 0.125), color=(17.0, 12.0, 4.0))]
    cam = dict(eye=(0.0, 1.0, -3.4), target=(0.0, 1.0, 0.0), up=(0.0, 1.0, 0.0),
               fovy=float(np.radians(40.0)), aspect=width / height)
    return RenderScene(verts, tris, tri_mat, materials, lights, cam, width, height, spp,
                       max_depth, light_sampler=0, seed=seed)


def doorway_scene(width=128, height=128, spp=16, max_depth=8, seed=1, res=16):
    """Two rooms joined by a doorway; the only light (a 0.3 x 0.3 quad of radiance 60) is in
    the far room, so everything the camera sees is lit indirectly through the opening."""
    g = 0
    parts = []

    def add(o, u, v, mat=g, n=res):
        parts.append((quad_mesh(o, u, v, n, n), mat))

    # room A (camera), z in [-1, 1]; inward normals
    add((-1, 0, -1), (0, 0, 2), (2, 0, 0))          # floor
    add((-1, 2, -1), (2, 0, 0), (0, 0, 2))          # ceiling
    add((-1, 0, -1), (0, 2, 0), (0, 0, 2), 1)       # left (red)
    add((1, 0, -1), (0, 0, 2), (0, 2, 0), 2)        # right (green)
    add((-1, 0, -1), (2, 0, 0), (0, 2, 0))          # front, behind the camera (+z)
    # partition: A side at z = 1 (-z normal), B side at z = 1.02 (+z normal), opening
    # x in [-0.25, 0.25], y in [0, 1]
    t = 0.02
    for (x0, x1, y0, y1) in [(-1, -0.25, 0, 2), (0.25, 1, 0, 2), (-0.25, 0.25, 1, 2)]:
        add((x0, y0, 1), (0, y1 - y0, 0), (x1 - x0, 0, 0), g, 8)
        add((x0, y0, 1 + t), (x1 - x0, 0, 0), (0, y1 - y0, 0), g, 8)
    # jambs and lintel of the opening
    add((-0.25, 0, 1), (0, 0, t), (0, 1, 0), g, 2)   # left jamb faces +x
    add((0.25, 0, 1), (0, 1, 0), (0, 0, t), g, 2)    # right jamb faces -x
    add((-0.25, 1, 1), (0.5, 0, 0), (0, 0, t), g, 2)  # lintel faces -y
    # room B, z in [1.02, 3]
    zb, lb = 1 + t, 3 - (1 + t)
    add((-1, 0, zb), (0, 0, lb), (2, 0, 0))         # floor
    add((-1, 2, zb), (2, 0, 0), (0, 0, lb))         # ceiling
    add((-1, 0, zb), (0, 2, 0), (0, 0, lb))         # left
    add((1, 0, zb), (0, 0, lb), (0, 2, 0))          # right
    add((-1, 0, 3), (0, 2, 0), (2, 0, 0))           # back
    verts, tris = concat(*[m for m, _ in parts])
    tri_mat = np.concatenate([np.full(len(m[1]), k, np.uint32) for m, k in parts])
    materials = [(0, (0.7, 0.7, 0.7), (1.0, 1.0)), (0, (0.65, 0.1, 0.1), (1.0, 1.0)),
                 (0, (0.12, 0.45, 0.15), (1.0, 1.0))]
    lights = [dict(kind=2, p0=(-0.15, 1.99, 2.4), p1=(0.15, 1.99, 2.4), p2=(-0.15, 1.99, 2.7),
                   color=(60.0, 55.0, 45.0))]
    cam = dict(eye=(0.3, 1.0, -0.9), target=(-0.1, 0.7, 1.0), up=(0.0, 1.0, 0.0),
               fovy=float(np.radians(60.0)), aspect=width / height)
    return RenderScene(verts, tris, tri_mat, materials, lights, cam, width, height, spp,
                       max_depth, light_sampler=0, seed=seed)


def rel_mse(img, ref, eps=1e-2):
    """mean over pixels and channels of (x - r)^2 / (r^2 + eps)."""
    img = np.asarray(img, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.mean((img - ref) ** 2 / (ref * ref + eps)))


def many_light_scene(width=1920, height=1080, spp=64, max_depth=8, rr_depth=3, grid=64,
                     face_res=4, n_lights=10_000, seed=5):
    """64 x 64 grid of boxes (4 x 4 quads per face -> 192 triangles per box, 786,432 in
    all) on a ground plane, lit by 10 K emissive triangles (lights only) with log-normal
    power (mu 0, sigma 1.5).  Materials alternate between rough conductors (alpha uniform
    0.05-0.8, F0 uniform) and coated dielectrics (F0 0.04, alpha uniform 0.05-0.8).
    Light-BVH sampler, max depth 8, Russian roulette from depth 3."""
    rng = np.random.default_rng(seed)
    meshes, mats, tri_mat = [], [], []
    half = grid / 2.0
    n_mat = 64
    for k in range(n_mat):
        a = float(rng.uniform(0.05, 0.8))
        if k % 2 == 0:
            mats.append((1, tuple(rng.uniform(0.2, 1.0, 3)), (a, a)))
        else:
            mats.append((0, tuple(rng.uniform(0.1, 0.9, 3)), (a, a)))
    ground = quad_mesh((-half, 0.0, -half), (0, 0, 2 * half), (2 * half, 0, 0), 64, 64)
    meshes.append(ground)
    tri_mat.append(np.zeros(len(ground[1]), np.uint32))
    for i in range(grid):
        for j in range(grid):
            sx, sz = rng.uniform(0.3, 0.8, 2)
            h = rng.uniform(0.2, 2.0)
            ang = rng.uniform(0, 90)
            pos = (-half + i + 0.5, 0.0, -half + j + 0.5)
            m = box_mesh(sx, h, sz, ang, pos, face_res)
            meshes.append(m)
            tri_mat.append(np.full(len(m[1]), rng.integers(0, n_mat), np.uint32))
    verts, tris = concat(*meshes)
    tri_mat = np.concatenate(tri_mat)
    # emissive triangles, facing down
    c = np.stack([rng.uniform(-half, half, n_lights), rng.uniform(2.5, 4.0, n_lights),
                  rng.uniform(-half, half, n_lights)], 1)
    power = np.exp(rng.normal(0.0, 1.5, n_lights))
    col = rng.uniform(0.5, 1.0, (n_lights, 3)) * power[:, None] * 4.0
    s = 0.08
    lights = [dict(kind=3, p0=c[k], p1=c[k] + (s, 0, 0), p2=c[k] + (0, 0, s), color=col[k])
              for k in range(n_lights)]
    cam = dict(eye=(0.0, 14.0, -34.0), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0),
               fovy=float(np.radians(45.0)), aspect=width / height)
    return RenderScene(verts, tris, tri_mat, mats, lights, cam, width, height, spp, max_depth,
                       light_sampler=2, rr_depth=rr_depth, seed=seed)


def guiding_scene(width=256, height=256, spp=16, max_depth=8, grid=64, face_res=4, seed=5,
                  radiance=(400.0, 380.0, 340.0)):
    """Learned-guiding workload: the many-light scene's box grid (GGX conductors and coated
    dielectrics on a ground plane) closed in by walls and a ceiling, lit only by one small
    quad light in a side chamber behind a 4 x 3 doorway, so every visible surface is lit
    indirectly through the opening.  The camera stands inside the main room."""
    rng = np.random.default_rng(seed)
    half = grid / 2.0
    H = 6.0
    meshes, tri_mat, mats = [], [], []
    n_mat = 16
    for k in range(n_mat):
        a = float(rng.uniform(0.05, 0.8))
        if k % 2 == 0:
            mats.append((1, tuple(rng.uniform(0.2, 1.0, 3)), (a, a)))
        else:
            mats.append((0, tuple(rng.uniform(0.1, 0.9, 3)), (a, a)))
    wall = len(mats)
    mats.append((0, (0.75, 0.75, 0.75), (1.0, 1.0)))

    def add(m, k):
        meshes.append(m)
        tri_mat.append(np.full(len(m[1]), k, np.uint32))

    # main room: floor, ceiling, walls with inward normals
    add(quad_mesh((-half, 0, -half), (0, 0, 2 * half), (2 * half, 0, 0), 32, 32), wall)
    add(quad_mesh((-half, H, -half), (2 * half, 0, 0), (0, 0, 2 * half), 32, 32), wall)
    add(quad_mesh((-half, 0, -half), (0, H, 0), (0, 0, 2 * half), 8, 32), wall)
    add(quad_mesh((half, 0, -half), (0, 0, 2 * half), (0, H, 0), 32, 8), wall)
    add(quad_mesh((-half, 0, -half), (2 * half, 0, 0), (0, H, 0), 32, 8), wall)
    # front wall at z = +half with a doorway x in [-2, 2], y in [0, 3]; both faces
    t = 0.2
    for (x0, x1, y0, y1) in [(-half, -2.0, 0, H), (2.0, half, 0, H), (-2.0, 2.0, 3.0, H)]:
        add(quad_mesh((x0, y0, half), (0, y1 - y0, 0), (x1 - x0, 0, 0), 8, 4), wall)
        add(quad_mesh((x0, y0, half + t), (x1 - x0, 0, 0), (0, y1 - y0, 0), 8, 4), wall)
    add(quad_mesh((-2.0, 0, half), (0, 0, t), (0, 3.0, 0), 1, 2), wall)
    add(quad_mesh((2.0, 0, half), (0, 3.0, 0), (0, 0, t), 2, 1), wall)
    add(quad_mesh((-2.0, 3.0, half), (4.0, 0, 0), (0, 0, t), 2, 1), wall)
    # side chamber z in [half + t, half + 8], x in [-8, 8]
    z0, z1, cx = half + t, half + 8.0, 8.0
    add(quad_mesh((-cx, 0, z0), (0, 0, z1 - z0), (2 * cx, 0, 0), 8, 8), wall)
    add(quad_mesh((-cx, H, z0), (2 * cx, 0, 0), (0, 0, z1 - z0), 8, 8), wall)
    add(quad_mesh((-cx, 0, z0), (0, H, 0), (0, 0, z1 - z0), 4, 8), wall)
    add(quad_mesh((cx, 0, z0), (0, 0, z1 - z0), (0, H, 0), 8, 4), wall)
    add(quad_mesh((-cx, 0, z1), (0, H, 0), (2 * cx, 0, 0), 4, 8), wall)
    # the box grid
    for i in range(grid):
        for j in range(grid):
            sx, sz = rng.uniform(0.3, 0.8, 2)
            h = rng.uniform(0.2, 2.0)
            ang = rng.uniform(0, 90)
            add(box_mesh(sx, h, sz, ang, (-half + i + 0.5, 0.0, -half + j + 0.5), face_res),
                int(rng.integers(0, n_mat)))
    verts, tris = concat(*meshes)
    tri_mat = np.concatenate(tri_mat)
    s = 1.0
    lights = [dict(kind=2, p0=(-s / 2, H - 0.01, z1 - 3.0), p1=(s / 2, H - 0.01, z1 - 3.0),
                   p2=(-s / 2, H - 0.01, z1 - 3.0 + s), color=radiance)]
    cam = dict(eye=(6.0, 4.5, half - 14.0), target=(-2.0, 0.5, half - 1.0), up=(0.0, 1.0, 0.0),
               fovy=float(np.radians(60.0)), aspect=width / height)
    return RenderScene(verts, tris, tri_mat, mats, lights, cam, width, height, spp, max_depth,
                       light_sampler=0, seed=seed)


_SPHERES = {}


def _unit_icosphere(subdiv):
    if subdiv not in _SPHERES:
        v, f = icosahedron()
        for _ in range(subdiv):
            v, f = subdivide(v, f)
        _SPHERES[subdiv] = (v, f)
    return _SPHERES[subdiv]


class BatchedScenes:
    """Many independent scenes for one tiled frame: scene k is a displaced icosphere plus
    random triangles (the traversal microbench's construction with seed `seed + k`), sized
    log-uniformly between `lo` and `hi` triangles.  The last `anim` fraction of every scene's
    triangles are random triangles with vertices of their own (a contiguous vertex range)
    that turn rigidly about the y axis from frame to frame.  All scenes share one quad light
    overhead and a grey Lambertian material; scene k is seen by camera k."""

    def __init__(self, n_scenes=32, lo=10_000, hi=2_000_000, seed=107, anim=0.25, tile=512):
        rng = np.random.default_rng(seed)
        self.sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), n_scenes)).astype(np.int64)
        self.meshes, self.anim_range = [], []
        for k, n in enumerate(self.sizes):
            s = 3
            while 20 * 4 ** (s + 1) <= (1.0 - anim) * n:
                s += 1
            v, f = _unit_icosphere(s)
            r = 1.0 + 0.1 * value_noise(v, seed + k)
            sphere = ((v * r[:, None]).astype(np.float32), f.astype(np.uint32))
            n_rand = max(int(n) - len(f), 1)
            rand = random_triangles(n_rand, seed + k)
            verts, tris = concat(sphere, rand)
            n_anim = max(1, int(round(anim * len(tris))))
            # the animated triangles are the last n_anim random ones: 3 vertices each, at the end
            self.anim_range.append((len(verts) - 3 * n_anim, 3 * n_anim))
            self.meshes.append((verts, tris))
        self.tile = tile
        self.cams = []
        for k in range(n_scenes):
            a = 2.0 * np.pi * k / n_scenes
            eye = (6.5 * np.sin(a), 2.5 + 0.5 * np.cos(3 * a), 6.5 * np.cos(a))
            self.cams.append(dict(eye=eye, target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0),
                                  fovy=float(np.radians(50.0)), aspect=1.0))
        self.materials = [(0, (0.7, 0.7, 0.7), (1.0, 1.0))]
        self.lights = [dict(kind=2, p0=(-1.0, 5.0, -1.0), p1=(1.0, 5.0, -1.0), p2=(-1.0, 5.0, 1.0),
                            color=(20.0, 19.0, 17.0))]
        self.tiles_x = 8 if n_scenes % 8 == 0 else n_scenes
        self.tiles_y = n_scenes // self.tiles_x

    @property
    def triangles(self):
        return int(sum(len(t) for _, t in self.meshes))

    @staticmethod
    def rotation(angle):
        """3 x 4 rigid transform: rotation about y by `angle` radians."""
        c, s = np.cos(angle), np.sin(angle)
        return np.array([[c, 0, s, 0], [0, 1, 0, 0], [-s, 0, c, 0]], np.float32)
