"""Scenes whose rendered value the mathematics fixes, for pinning the integrator
(Alpine::traceRayBatch / estimateDirectIllumination, alpine.cpp:290-432) on the oracle and on
the GPU path tracer.

All materials are the reference Dielectric with alpha = 1, i.e. a pure Lambertian of albedo
rho (dielectric.cpp:29-45: the GGX coat only exists for alpha^2 < 1), whose sampling
estimator is rho (lambertian.cpp:18).  Lights are not geometry: camera rays pass through them.
"""
import numpy as np

from alpine_amd import scenes

LAMBERT = 0  # APG_MAT_DIELECTRIC with alpha (1, 1)
DISK, POINT, QUAD, TRIANGLE = 0, 1, 2, 3
LS_UNIFORM = 0


def plane(size=100.0):
    """The square |x|, |z| <= size at y = 0, geometric normal +y."""
    return scenes.quad_mesh((-size, 0.0, -size), (0.0, 0.0, 2 * size), (2 * size, 0.0, 0.0), 1, 1)


def down_camera(height, fovy, target=(0.0, 0.0, 0.0)):
    return dict(eye=(target[0], height, target[2]), target=target, up=(0.0, 0.0, 1.0), fovy=fovy,
                aspect=1.0)


class Case:
    """A render setup: geometry, one material, lights, camera, optional shading normals."""

    def __init__(self, verts, tris, rho, lights, camera, max_depth, width=64, height=64, spp=16,
                 background=(0.0, 0.0, 0.0), normals=None, rr_depth=0, seed=1, extra=None):
        self.verts, self.tris = verts, tris
        self.rho = np.asarray(rho, np.float64)
        self.lights, self.camera = lights, camera
        self.max_depth, self.width, self.height, self.spp = max_depth, width, height, spp
        self.background = background
        self.normals = normals          # per-vertex shading normals or None (ns = ng)
        self.rr_depth, self.seed = rr_depth, seed
        self.tri_mat = np.zeros(len(tris), np.uint32)
        if extra is not None:           # a second mesh of the same material (occluders)
            ev, et = extra
            self.tris = np.concatenate([self.tris, et + len(self.verts)])
            self.verts = np.concatenate([self.verts, ev])
            if self.normals is not None:
                en = np.tile(np.array([[0.0, 1.0, 0.0]], np.float32), (len(ev), 1))
                self.normals = np.concatenate([self.normals, en])
            self.tri_mat = np.zeros(len(self.tris), np.uint32)

    @property
    def material(self):
        return (LAMBERT, tuple(float(x) for x in self.rho), (1.0, 1.0))


def tilted_normals(verts, theta):
    """Every vertex's shading normal tilted by theta from +y towards +x."""
    n = np.array([np.sin(theta), np.cos(theta), 0.0], np.float32)
    return np.tile(n, (len(verts), 1))


def polygon_irradiance(pts, p, n):
    """Irradiance at p (normal n) from a polygon of unit radiance seen entirely above the
    surface: E = |sum_i theta_i (u_i x u_{i+1}) / |u_i x u_{i+1}| / 2 . n| (Lambert's formula,
    u_i the unit vectors from p to the vertices, theta_i the angle between them)."""
    u = [np.asarray(q, np.float64) - np.asarray(p, np.float64) for q in pts]
    u = [x / np.linalg.norm(x) for x in u]
    phi = np.zeros(3)
    for i in range(len(u)):
        a, b = u[i], u[(i + 1) % len(u)]
        c = np.cross(a, b)
        phi += 0.5 * np.arccos(np.clip(a @ b, -1.0, 1.0)) * c / np.linalg.norm(c)
    return abs(phi @ np.asarray(n, np.float64))


def inward_sphere(subdiv=5, radius=1.0):
    """Icosphere with every triangle facing the centre; returns (verts, tris, h_min) with
    h_min the smallest distance from the centre to a triangle's plane."""
    v, f = scenes.icosahedron()
    for _ in range(subdiv):
        v, f = scenes.subdivide(v, f)
    v = v / np.linalg.norm(v, axis=1, keepdims=True) * radius
    f = f[:, ::-1].copy()  # outward -> inward winding
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    n = np.cross(b - a, c - a)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    h = -(n * a).sum(axis=1)  # centre is on the side the normal points to
    assert np.all(h > 0), "triangles face the centre"
    return v.astype(np.float32), f.astype(np.uint32), float(h.min())


# ---- the cases -------------------------------------------------------------------------
BG = (0.9, 1.3, 2.0)
RHO = (0.73, 0.5, 0.25)


def open_plane(max_depth, theta=0.0, spp=16):
    """Plane under a uniform background, no lights: after one bounce every path escapes, so a
    pixel is rho * B * P(the cosine-sampled continuation about ns lies above the geometric
    plane), P = (1 + cos theta) / 2 for shading normals tilted by theta (paths with
    dot(wi, ng) <= 0 are killed, alpine.cpp:350-355); 0 when max_depth = 1 (the continuation
    is never traced)."""
    v, f = plane()
    normals = tilted_normals(v, theta) if theta else None
    c = Case(v, f, RHO, [], down_camera(1.0, np.radians(60.0)), max_depth, spp=spp, background=BG,
             normals=normals)
    p = 0.0 if max_depth < 2 else 0.5 * (1.0 + np.cos(theta))
    c.expect = np.asarray(RHO) * np.asarray(BG) * p
    c.bernoulli = 0.0 < p < 1.0
    return c


def disk_on_axis(theta=0.0, spp=64, h=1.0, r=0.5, power=40.0, color=(1.0, 0.8, 0.6)):
    """One-sided disk light of radius r facing the plane at height h, seen by a narrow camera
    at its axis point: E = L pi sin^2(a) cos(theta), sin^2(a) = r^2 / (r^2 + h^2), L = P /
    (pi r^2 pi) (disk_light.cpp:9-28), pixel = rho E / pi (theta: shading-normal tilt, the
    whole cone inside the tilted hemisphere)."""
    v, f = plane()
    light = dict(kind=DISK, p0=(0.0, h, 0.0), normal=(0.0, -1.0, 0.0), color=color, power=power, radius=r)
    normals = tilted_normals(v, theta) if theta else None
    c = Case(v, f, RHO, [light], down_camera(3.0, np.radians(0.2)), 1, spp=spp, normals=normals)
    L = np.asarray(color) * power / (np.pi * r * r * np.pi)
    c.expect = np.asarray(RHO) * L * (r * r / (r * r + h * h)) * np.cos(theta)
    return c


def point_above(h=1.5, power=30.0, color=(0.4, 1.0, 0.7)):
    """Point light of power P at height h on the camera's axis: intensity P / (4 pi)
    (point_light.cpp:15-20), pixel = rho / pi * I / h^2, the same for every sample."""
    v, f = plane()
    light = dict(kind=POINT, p0=(0.0, h, 0.0), color=color, power=power)
    c = Case(v, f, RHO, [light], down_camera(3.0, np.radians(0.2)), 1, spp=4)
    I = np.asarray(color) * power / (4.0 * np.pi)
    c.expect = np.asarray(RHO) / np.pi * I / (h * h)
    return c


def polygon_above(kind, spp=64, radiance=(2.0, 1.5, 1.0)):
    """A quad or triangle light facing the plane at height 1, off the camera's axis:
    pixel = rho / pi * L * E(polygon) by Lambert's formula."""
    p0, p1, p2 = (-0.3, 1.0, -0.2), (0.5, 1.0, -0.2), (-0.3, 1.0, 0.4)  # e0 x e1 points down
    v, f = plane()
    light = dict(kind=kind, p0=p0, p1=p1, p2=p2, color=radiance)
    c = Case(v, f, RHO, [light], down_camera(3.0, np.radians(0.2)), 1, spp=spp)
    if kind == QUAD:
        pts = [p0, p1, np.add(p1, np.subtract(p2, p0)), p2]
    else:
        pts = [p0, p1, p2]
    c.expect = np.asarray(RHO) / np.pi * np.asarray(radiance) * polygon_irradiance(pts, (0, 0, 0), (0, 1, 0))
    return c


def grazing_point_with_ledge(theta=np.radians(80.0)):
    """Shading normals tilted 80 degrees towards a point light low on the horizon, and a thin
    horizontal ledge at height 8e-4 between 0.2 and 0.5 along x.  The reference offsets
    shadow rays by 1e-3 along the shading normal (alpine.cpp:389-394): the ray starts 1.7e-4
    above the plane, climbs 0.002 per unit and meets the ledge near x = 0.31, so every pixel
    is 0.  Offset along ng (1e-3 up) the ray passes above the ledge and the pixel would be
    rho / pi * I cos_s / d^2."""
    v, f = plane()
    ledge = scenes.quad_mesh((0.2, 8e-4, -1.0), (0.0, 0.0, 2.0), (0.3, 0.0, 0.0), 1, 1)
    light = dict(kind=POINT, p0=(10.0, 0.02, 0.0), color=(1.0, 1.0, 1.0), power=400.0)
    c = Case(v, f, RHO, [light], down_camera(3.0, np.radians(0.2)), 1, spp=2,
             normals=tilted_normals(v, theta), extra=ledge)
    c.expect = np.zeros(3)
    d = np.array([10.0, 0.02, 0.0])
    ns = np.array([np.sin(theta), np.cos(theta), 0.0])
    c.unshadowed = np.asarray(RHO) / np.pi * 400.0 / (4 * np.pi) * (ns @ d / np.linalg.norm(d)) / (d @ d)
    return c


def furnace(max_depth, rho, rr_depth=0, spp=16, subdiv=5, power=20.0):
    """Closed sphere of radius 1 facing inwards, point light and camera at its centre.  Every
    path has max_depth vertices (closed, the Lambert continuation never fails); at each the
    light is unoccluded with cos = h_t / d, so the direct term is rho I h_t / (pi d^3), between
    rho I h_min / pi (d <= 1) and rho I / (pi h_min^2) (d >= h_t).  A sample is
    sum_k rho^k * direct_k, so every pixel lies in S * [h_min, 1 / h_min^2] * rho I / pi with
    S = (1 - rho^D) / (1 - rho) (exactly, without Russian roulette)."""
    v, f, hmin = inward_sphere(subdiv)
    light = dict(kind=POINT, p0=(0.0, 0.0, 0.0), color=(1.0, 1.0, 1.0), power=power)
    cam = dict(eye=(0.0, 0.0, 0.0), target=(0.3, 0.2, 1.0), up=(0.0, 1.0, 0.0), fovy=np.radians(90.0), aspect=1.0)
    rr = np.asarray(rho, np.float64)
    c = Case(v, f, rho, [light], cam, max_depth, spp=spp, rr_depth=rr_depth)
    S = (1.0 - rr ** max_depth) / (1.0 - rr)
    base = rr * (power / (4.0 * np.pi)) / np.pi
    c.lo, c.hi = S * base * hmin, S * base / (hmin * hmin)
    c.expect = 0.5 * (c.lo + c.hi)
    return c


# ---- running a case --------------------------------------------------------------------
def oracle_image(c, mis_mode=0, plant=0, spp=None, nthreads=8):
    """Per-pixel mean radiance (h, w, 3) of the oracle integrator."""
    import oracle
    mats = [oracle.make_material(*c.material)]
    lights = [oracle.make_light(**l) for l in c.lights]
    cam = oracle.make_camera(**c.camera)
    s = oracle.OracleScene(c.verts, c.tris, c.tri_mat, mats, lights, LS_UNIFORM, cam, c.background)
    if c.normals is not None:
        s.set_shading(normals=c.normals)
    n = spp or c.spp
    states = oracle.sampler_states(c.width * c.height, c.seed)
    _, acc = s.render(c.width, c.height, n, c.max_depth, states, None, c.rr_depth, nthreads,
                      mis_mode=mis_mode, plant=plant)
    return acc / n


def gpu_image(c, spp=None, slots=1, rr_depth=None, watertight=False):
    """Per-pixel mean radiance (h, w, 3) of the GPU path tracer (FP64 accumulation);
    watertight: trace with the watertight triangle test instead of the reference's."""
    import alpine_amd as aa
    acc = aa.GpuAccelerator(0)
    sid = acc.append_mesh(c.verts, c.tris)
    acc.update_scene()
    if watertight:
        acc.set_triangle_test(aa.TRI_WATERTIGHT)
    if c.normals is not None:
        acc.set_shading(sid, normals=c.normals)
    acc.set_materials([aa.material(*c.material)], c.tri_mat)
    acc.set_lights([aa.light(**l) for l in c.lights], LS_UNIFORM)
    acc.set_camera(aa.camera(**c.camera))
    n = spp or c.spp
    acc.render_reset(c.width, c.height, c.seed, slots=slots)
    acc.render(n, c.max_depth, c.rr_depth if rr_depth is None else rr_depth, c.background, fp64=True)
    _, r64, total = acc.render_read(fp64=True)
    acc.close()
    assert total == n
    return r64 / n
