"""ORACLE -- test infrastructure only.

ctypes bindings for the CPU checker:

* ``liboracle.so`` -- the plain-C restatement of the reference's binary BVH
  (oracle/bvh.c, following alpine src/accelerators/bvh.cpp and
  bvh_common.cpp);
* ``_ref/libalpine_ref.so`` -- the reference's own bvh.cpp/bvh_common.cpp
  compiled here by oracle/Makefile, behind oracle/ref_harness.cpp.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package.  The product (alpine_amd) never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libalpine_ref.so")
REF_SRC = "/root/reference"

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u16p = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


def build(ref: bool | None = None) -> None:
    """Compile the oracle (and the reference harness when the reference tree is present)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if ref is None:
        ref = os.path.isdir(os.path.join(REF_SRC, "src"))
    if ref:
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)
        # the drop-in adapter beside the reference's Bvh / samplers (needs alpine_amd/libapg.so)
        if os.path.exists(os.path.join(HERE, "..", "alpine_amd", "libapg.so")):
            subprocess.run(["make", "-s", "-C", HERE, "adapter"], check=True)
        # and its dummy form (no GPU library), the build without ALPINE_ENABLE_HIP
        subprocess.run(["make", "-s", "-C", HERE, "dummy"], check=True)


class _Stats(C.Structure):
    _fields_ = [
        ("node_count", C.c_uint32),
        ("leaf_count", C.c_uint32),
        ("max_depth", C.c_uint32),
        ("max_leaf_size", C.c_uint32),
        ("sah_cost", C.c_double),
        ("sah_cost_leaf", C.c_double),
    ]


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build(ref=False)
        L = C.CDLL(LIB_PATH)
        L.oracle_bvh_create.restype = C.c_void_p
        L.oracle_bvh_create.argtypes = [_f32p, C.c_uint32, _u32p, C.c_uint32, C.c_uint32]
        L.oracle_wide_create.restype = C.c_void_p
        L.oracle_wide_create.argtypes = [_f32p, C.c_uint32, _u32p, C.c_uint32]
        L.oracle_wide_destroy.argtypes = [C.c_void_p]
        L.oracle_wide_node_count.restype = C.c_uint32
        L.oracle_wide_node_count.argtypes = [C.c_void_p]
        L.oracle_wide_intersect.argtypes = [C.c_void_p, _f32p, _f32p, C.c_uint32, _u32p, _f32p,
                                            _f32p, _f32p, C.c_uint32]
        L.oracle_wide_occluded.argtypes = [C.c_void_p, _f32p, _f32p, C.c_void_p, C.c_uint32,
                                           _u8p, C.c_uint32]
        L.oracle_wide_brute.argtypes = [C.c_void_p, _f32p, _f32p, C.c_uint32, _u32p, _f32p]
        L.oracle_bvh_destroy.argtypes = [C.c_void_p]
        L.oracle_bvh_intersect.argtypes = [C.c_void_p, _f32p, _f32p, C.c_uint32, _u32p, _f32p,
                                           _f32p, _f32p, C.c_uint32]
        L.oracle_bvh_occluded.argtypes = [C.c_void_p, _f32p, _f32p, C.c_void_p, C.c_uint32,
                                          _u8p, C.c_uint32]
        L.oracle_brute_intersect.argtypes = [C.c_void_p, _f32p, _f32p, C.c_uint32, _u32p,
                                             _f32p, _f32p, _f32p]
        L.oracle_brute_occluded.argtypes = [C.c_void_p, _f32p, _f32p, C.c_void_p, C.c_uint32,
                                            _u8p]
        L.oracle_triangle_intersect.restype = C.c_int
        L.oracle_triangle_intersect.argtypes = [C.c_void_p, C.c_uint32, _f32p, _f32p,
                                                C.POINTER(C.c_float), C.POINTER(C.c_float),
                                                C.POINTER(C.c_float)]
        L.oracle_bvh_get_stats.argtypes = [C.c_void_p, C.POINTER(_Stats)]
        L.oracle_bvh_leaf_keys.argtypes = [C.c_void_p, _u32p]
        L.oracle_bvh_ordered_prims.argtypes = [C.c_void_p, _u32p]
        L.oracle_bvh_nodes.argtypes = [C.c_void_p, _f32p, _u32p, _u16p, _u8p]
        L.oracle_find_split.restype = C.c_int
        L.oracle_find_split.argtypes = [_f32p, C.c_uint32, C.POINTER(C.c_uint32),
                                        C.POINTER(C.c_uint32), C.POINTER(C.c_float)]
        L.oracle_surface_area.restype = C.c_float
        L.oracle_surface_area.argtypes = [_f32p]
        L.oracle_box_intersect.restype = C.c_int
        L.oracle_box_intersect.argtypes = [_f32p, _f32p, _f32p, C.c_float]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_LIB_PATH)


def ref_lib():
    global _ref
    if _ref is None:
        L = C.CDLL(REF_LIB_PATH)
        L.ref_bvh_create.restype = C.c_void_p
        L.ref_bvh_create.argtypes = [_f32p, C.c_uint32, _u32p, C.c_uint32]
        L.ref_bvh_destroy.argtypes = [C.c_void_p]
        L.ref_bvh_intersect.argtypes = [C.c_void_p, _f32p, _f32p, C.c_uint32, _u32p, _f32p,
                                        _f32p, _f32p, C.c_uint32]
        L.ref_bvh_occluded.argtypes = [C.c_void_p, _f32p, _f32p, C.c_void_p, C.c_uint32, _u8p,
                                       C.c_uint32]
        L.ref_tree_stats.restype = C.c_int
        L.ref_tree_stats.argtypes = [_f32p, C.c_uint32, _u32p, C.c_uint32, C.c_uint32,
                                     C.POINTER(C.c_double), C.POINTER(C.c_double),
                                     C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                     C.POINTER(C.c_uint32), _u32p]
        _ref = L
    return _ref


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _tfar_arg(tfar, n):
    if tfar is None:
        return None, None
    arr = _f32(np.broadcast_to(np.asarray(tfar, np.float32), (n,)))
    return arr, arr.ctypes.data_as(C.c_void_p)


class Hits:
    """Closest-hit result arrays (prim = 0xFFFFFFFF on a miss)."""

    def __init__(self, n):
        self.prim = np.empty(n, np.uint32)
        self.t = np.empty(n, np.float32)
        self.b0 = np.empty(n, np.float32)
        self.b1 = np.empty(n, np.float32)


class OracleWideBvh:
    """The reference WideBvh (wide_bvh.cpp) with the ISPC kernels' arithmetic, restated; the
    8-wide CPU path timed beside the GPU.  Parity against the reference is unpinned (ISPC is
    absent); brute_intersect uses the same triangle test and pins the traversal."""

    def __init__(self, verts, tris):
        self.verts = _f32(verts).reshape(-1, 3)
        self.tris = _u32(tris).reshape(-1, 3)
        self._h = lib().oracle_wide_create(self.verts.reshape(-1), len(self.verts),
                                           self.tris.reshape(-1), len(self.tris))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().oracle_wide_destroy(h)
            self._h = None

    @property
    def node_count(self):
        return lib().oracle_wide_node_count(self._h)

    def intersect(self, org, dir, nthreads: int = 1) -> "Hits":
        org, dir = _f32(org), _f32(dir)
        n = len(org)
        h = Hits(n)
        lib().oracle_wide_intersect(self._h, org.reshape(-1), dir.reshape(-1), n, h.prim, h.t,
                                    h.b0, h.b1, nthreads)
        return h

    def occluded(self, org, dir, tfar=None, nthreads: int = 1):
        org, dir = _f32(org), _f32(dir)
        n = len(org)
        out = np.empty(n, np.uint8)
        keep, ptr = _tfar_arg(tfar, n)
        lib().oracle_wide_occluded(self._h, org.reshape(-1), dir.reshape(-1), ptr, n, out, nthreads)
        del keep
        return out

    def brute_intersect(self, org, dir):
        org, dir = _f32(org), _f32(dir)
        n = len(org)
        prim = np.empty(n, np.uint32)
        t = np.empty(n, np.float32)
        lib().oracle_wide_brute(self._h, org.reshape(-1), dir.reshape(-1), n, prim, t)
        return prim, t


class OracleBvh:
    """The reference binary BVH, restated (bvh.cpp:60-319)."""

    def __init__(self, verts, tris, leaf_threshold: int = 8):
        self.verts = _f32(verts).reshape(-1, 3)
        self.tris = _u32(tris).reshape(-1, 3)
        self.nt = len(self.tris)
        self._h = lib().oracle_bvh_create(self.verts.reshape(-1), len(self.verts),
                                          self.tris.reshape(-1), self.nt, leaf_threshold)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().oracle_bvh_destroy(h)
            self._h = None

    def intersect(self, org, dir, nthreads: int = 1) -> Hits:
        org, dir = _f32(org), _f32(dir)
        n = len(org)
        h = Hits(n)
        lib().oracle_bvh_intersect(self._h, org.reshape(-1), dir.reshape(-1), n, h.prim, h.t,
                                   h.b0, h.b1, nthreads)
        return h

    def occluded(self, org, dir, tfar=None, nthreads: int = 1):
        org, dir = _f32(org), _f32(dir)
        n = len(org)
        out = np.empty(n, np.uint8)
        keep, ptr = _tfar_arg(tfar, n)
        lib().oracle_bvh_occluded(self._h, org.reshape(-1), dir.reshape(-1), ptr, n, out,
                                  nthreads)
        del keep
        return out

    def brute_intersect(self, org, dir) -> Hits:
        org, dir = _f32(org), _f32(dir)
        n = len(org)
        h = Hits(n)
        lib().oracle_brute_intersect(self._h, org.reshape(-1), dir.reshape(-1), n, h.prim, h.t,
                                     h.b0, h.b1)
        return h

    def brute_occluded(self, org, dir, tfar=None):
        org, dir = _f32(org), _f32(dir)
        n = len(org)
        out = np.empty(n, np.uint8)
        keep, ptr = _tfar_arg(tfar, n)
        lib().oracle_brute_occluded(self._h, org.reshape(-1), dir.reshape(-1), ptr, n, out)
        del keep
        return out

    def triangle(self, tri: int, org, dir):
        t, b0, b1 = C.c_float(), C.c_float(), C.c_float()
        hit = lib().oracle_triangle_intersect(self._h, tri, _f32(org), _f32(dir), C.byref(t),
                                              C.byref(b0), C.byref(b1))
        return bool(hit), t.value, b0.value, b1.value

    def stats(self) -> dict:
        s = _Stats()
        lib().oracle_bvh_get_stats(self._h, C.byref(s))
        return {k: getattr(s, k) for k, _ in _Stats._fields_}

    def leaf_keys(self):
        out = np.empty(self.nt, np.uint32)
        lib().oracle_bvh_leaf_keys(self._h, out)
        return out

    def ordered_prims(self):
        out = np.empty(self.nt, np.uint32)
        lib().oracle_bvh_ordered_prims(self._h, out)
        return out

    def nodes(self):
        n = self.stats()["node_count"]
        bbox = np.empty((n, 6), np.float32)
        off = np.empty(n, np.uint32)
        cnt = np.empty(n, np.uint16)
        dim = np.empty(n, np.uint8)
        lib().oracle_bvh_nodes(self._h, bbox.reshape(-1), off, cnt, dim)
        return bbox, off, cnt, dim


def find_split(boxes6):
    b = _f32(boxes6).reshape(-1)
    d, s, c = C.c_uint32(), C.c_uint32(), C.c_float()
    ok = lib().oracle_find_split(b, len(b) // 6, C.byref(d), C.byref(s), C.byref(c))
    return bool(ok), d.value, s.value, c.value


def surface_area(box6):
    return lib().oracle_surface_area(_f32(box6))


def box_intersect(box6, org, dir, tfar):
    return bool(lib().oracle_box_intersect(_f32(box6), _f32(org), _f32(dir), tfar))


class RefBvh:
    """The reference's compiled Bvh (oracle/_ref); only where the reference tree was present."""

    def __init__(self, verts, tris):
        self.verts = _f32(verts).reshape(-1, 3)
        self.tris = _u32(tris).reshape(-1, 3)
        self._h = ref_lib().ref_bvh_create(self.verts.reshape(-1), len(self.verts),
                                           self.tris.reshape(-1), len(self.tris))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            ref_lib().ref_bvh_destroy(h)
            self._h = None

    def intersect(self, org, dir, nthreads: int = 1) -> Hits:
        org, dir = _f32(org), _f32(dir)
        n = len(org)
        h = Hits(n)
        ref_lib().ref_bvh_intersect(self._h, org.reshape(-1), dir.reshape(-1), n, h.prim, h.t,
                                    h.b0, h.b1, nthreads)
        return h

    def occluded(self, org, dir, tfar=None, nthreads: int = 1):
        org, dir = _f32(org), _f32(dir)
        n = len(org)
        out = np.empty(n, np.uint8)
        keep, ptr = _tfar_arg(tfar, n)
        ref_lib().ref_bvh_occluded(self._h, org.reshape(-1), dir.reshape(-1), ptr, n, out,
                                   nthreads)
        del keep
        return out


def ref_tree_stats(verts, tris, leaf_threshold: int = 8) -> dict:
    v = _f32(verts).reshape(-1)
    t = _u32(tris).reshape(-1)
    nt = len(t) // 3
    keys = np.empty(nt, np.uint32)
    sah, sahl = C.c_double(), C.c_double()
    nc, lc, md = C.c_uint32(), C.c_uint32(), C.c_uint32()
    ref_lib().ref_tree_stats(v, len(v) // 3, t, nt, leaf_threshold, C.byref(sah), C.byref(sahl),
                             C.byref(nc), C.byref(lc), C.byref(md), keys)
    return {"sah_cost": sah.value, "sah_cost_leaf": sahl.value, "node_count": nc.value,
            "leaf_count": lc.value, "max_depth": md.value, "leaf_keys": keys}


# ---------------------------------------------------------------------------
# Render path (oracle/render.c) and the reference's render TUs (oracle/_ref)
# ---------------------------------------------------------------------------
REF_RENDER_LIB_PATH = os.path.join(HERE, "_ref", "libalpine_ref_render.so")

MAT_DIELECTRIC, MAT_CONDUCTOR = 0, 1
LIGHT_DISK, LIGHT_POINT, LIGHT_QUAD, LIGHT_TRIANGLE, LIGHT_POINT_INTENSITY = 0, 1, 2, 3, 4
LS_UNIFORM, LS_POWER, LS_BVH = 0, 1, 2


class Material(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("base_color", C.c_float * 3), ("alpha", C.c_float * 2),
                ("base_color_tex", C.c_uint32), ("normal_tex", C.c_uint32)]


class Shading(C.Structure):
    _fields_ = [("tri_normals", C.c_void_p), ("tri_tangents", C.c_void_p), ("tri_uvs", C.c_void_p),
                ("ntargets", C.c_uint32), ("tri_dnormals", C.c_void_p),
                ("tri_dtangents", C.c_void_p), ("weights", C.c_void_p)]


class Light(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("p0", C.c_float * 3), ("p1", C.c_float * 3),
                ("p2", C.c_float * 3), ("normal", C.c_float * 3), ("color", C.c_float * 3),
                ("power", C.c_float), ("radius", C.c_float)]


class Camera(C.Structure):
    _fields_ = [("eye", C.c_float * 3), ("target", C.c_float * 3), ("up", C.c_float * 3),
                ("fovy", C.c_float), ("aspect", C.c_float)]


class RenderParams(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32),
                ("max_depth", C.c_uint32), ("rr_depth", C.c_uint32), ("mis_mode", C.c_uint32),
                ("plant", C.c_uint32), ("crop_x", C.c_uint32), ("crop_y", C.c_uint32),
                ("crop_w", C.c_uint32), ("crop_h", C.c_uint32)]


# Test switches of the integrator (oracle/render.h): one direct-light strategy only, and
# planted mistakes that the closed-form tests must catch.
MIS_BOTH, MIS_LIGHT_ONLY, MIS_BSDF_ONLY = 0, 1, 2
PLANT_MIS_WEIGHT, PLANT_NO_NG_KILL, PLANT_SHADOW_NG, PLANT_NO_THROUGHPUT = 1, 2, 4, 8


def _render_sigs(L):
    vp = C.c_void_p
    L.orc_scene_create.restype = vp
    L.orc_scene_create.argtypes = [_f32p, C.c_uint32, _u32p, C.c_uint32, _u32p, vp, C.c_uint32,
                                   vp, vp, C.c_uint32, C.c_int, vp, vp]
    L.orc_scene_destroy.argtypes = [vp]
    L.orc_render.argtypes = [vp, vp, _u32p, np.ctypeslib.ndpointer(np.float64, flags="C"),
                             vp, C.c_uint32]
    L.orc_scene_set_shading.argtypes = [vp, C.POINTER(Shading)]
    L.orc_scene_set_textures.argtypes = [vp, vp, vp, C.c_uint32]
    L.orc_scene_set_environment.argtypes = [vp, vp, C.c_uint32, C.c_uint32]
    L.orc_shading_attributes.argtypes = [vp, C.c_uint32, C.c_float, C.c_float] + [_f32p] * 5
    L.orc_environment_sample.argtypes = [vp, _f32p, _f32p]
    L.orc_sampler_reset.argtypes = [_u32p, C.c_uint32]
    L.orc_sampler_get1d.restype = C.c_float
    L.orc_sampler_get1d.argtypes = [_u32p]
    L.orc_camera_ray.argtypes = [vp, C.c_float, C.c_float, _f32p, _f32p]
    L.orc_mat_sample.restype = C.c_int
    L.orc_mat_sample.argtypes = [vp, _f32p, _f32p, _f32p, _f32p, C.POINTER(C.c_float)]
    L.orc_mat_eval.argtypes = [vp, _f32p, _f32p, _f32p, C.POINTER(C.c_float)]
    L.orc_light_sampler_sample.restype = C.c_int
    L.orc_light_sampler_sample.argtypes = [vp, C.c_float, _f32p, _f32p, C.POINTER(C.c_float)]
    L.orc_light_sampler_pdf.restype = C.c_float
    L.orc_light_sampler_pdf.argtypes = [vp, C.c_int, _f32p, _f32p]
    L.orc_light_sample.restype = C.c_int
    L.orc_light_sample.argtypes = [vp, C.c_int, _f32p, _f32p, _f32p, _f32p,
                                   C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.orc_light_pdf.restype = C.c_float
    L.orc_light_pdf.argtypes = [vp, C.c_int, _f32p, _f32p, C.POINTER(C.c_float)]
    L.orc_postprocess.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.c_int,
                                  np.ctypeslib.ndpointer(np.uint8, flags="C")]


_render_ready = False


def rlib():
    global _render_ready
    L = lib()
    if not _render_ready:
        _render_sigs(L)
        _render_ready = True
    return L


def make_material(kind, base_color, alpha, base_color_tex=None, normal_tex=None):
    return Material(kind, (C.c_float * 3)(*base_color), (C.c_float * 2)(*alpha),
                    0 if base_color_tex is None else int(base_color_tex) + 1,
                    0 if normal_tex is None else int(normal_tex) + 1)


def make_light(kind, p0=(0, 0, 0), p1=(0, 0, 0), p2=(0, 0, 0), normal=(0, 1, 0),
               color=(1, 1, 1), power=1.0, radius=1.0):
    f3 = lambda v: (C.c_float * 3)(*[float(x) for x in v])
    return Light(kind, f3(p0), f3(p1), f3(p2), f3(normal), f3(color), float(power), float(radius))


def make_camera(eye, target, up, fovy, aspect):
    f3 = lambda v: (C.c_float * 3)(*[float(x) for x in v])
    return Camera(f3(eye), f3(target), f3(up), float(fovy), float(aspect))


class OracleScene:
    """The reference integrator restated (alpine.cpp:290-432) over the oracle BVH."""

    def __init__(self, verts, tris, tri_mat, materials, lights, light_sampler=LS_BVH,
                 camera=None, background=(0.0, 0.0, 0.0), normals=None):
        L = rlib()
        self.verts = _f32(verts).reshape(-1, 3)
        self.tris = _u32(tris).reshape(-1, 3)
        self.tri_mat = _u32(tri_mat)
        self._mats = (Material * max(1, len(materials)))(*materials)
        self._lights = (Light * max(1, len(lights)))(*lights)
        self._cam = camera
        self._bg = (C.c_float * 3)(*background)
        self._normals = None if normals is None else _f32(normals).reshape(-1)
        self._h = L.orc_scene_create(
            self.verts.reshape(-1), len(self.verts), self.tris.reshape(-1), len(self.tris),
            self.tri_mat, C.cast(self._mats, C.c_void_p), len(materials),
            None if self._normals is None else self._normals.ctypes.data_as(C.c_void_p),
            C.cast(self._lights, C.c_void_p), len(lights), light_sampler,
            C.cast(C.pointer(camera), C.c_void_p) if camera is not None else None,
            C.cast(self._bg, C.c_void_p))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            rlib().orc_scene_destroy(h)
            self._h = None

    def render(self, width, height, spp, max_depth, states=None, accum=None, rr_depth=0,
               nthreads=8, aov=None, mis_mode=MIS_BOTH, plant=0, crop=None):
        """aov: optional (height, width, 6) float64 array receiving the first hit's base colour
        and shading normal sums.  crop = (x0, y0, w, h): trace only that window of the frame;
        states and accum then cover the window (states seeded by frame pixel index, see
        crop_states)."""
        cx, cy, cw, ch = crop if crop is not None else (0, 0, 0, 0)
        pw, ph = (cw, ch) if crop is not None else (width, height)
        if crop is not None:
            assert cw > 0 and ch > 0 and cx + cw <= width and cy + ch <= height
        if states is None:
            states = sampler_states(pw * ph) if crop is None else crop_states(width, crop)
        assert states.shape == (pw * ph, 4)
        if accum is None:
            accum = np.zeros((ph, pw, 3), np.float64)
        assert accum.size == pw * ph * 3
        if aov is not None:
            assert aov.dtype == np.float64 and aov.flags.c_contiguous and aov.size == pw * ph * 6
        rp = RenderParams(width, height, spp, max_depth, rr_depth, mis_mode, plant, cx, cy, cw, ch)
        rlib().orc_render(self._h, C.byref(rp), states.reshape(-1), accum.reshape(-1),
                          None if aov is None else aov.ctypes.data_as(C.c_void_p), nthreads)
        return states, accum

    def set_shading(self, normals=None, normal_idx=None, tangents=None, uvs=None, uv_idx=None,
                    dnormals=None, dtangents=None, weights=None):
        """Per-vertex attributes (indexed like the GPU API: normal_idx / uv_idx default to the
        position indices), expanded per triangle for the restatement.  dnormals / dtangents:
        (ntargets, nn, 3) morph deltas of the normals / tangents, weights (ntargets)."""
        t = self.tris
        ni = t if normal_idx is None else _u32(normal_idx).reshape(-1, 3)
        ui = t if uv_idx is None else _u32(uv_idx).reshape(-1, 3)
        keep = []

        def exp(a, idx, cols):
            if a is None:
                return None
            e = np.ascontiguousarray(_f32(a).reshape(-1, cols)[idx].reshape(-1))
            keep.append(e)
            return e.ctypes.data_as(C.c_void_p)

        nt = 0
        dn = dt = w = None
        if normals is not None and weights is not None:
            w = _f32(weights).reshape(-1)
            nt = len(w)
            nn = len(_f32(normals).reshape(-1, 3))
            z = np.zeros((nt, nn, 3), np.float32)
            dn_a = z if dnormals is None else _f32(dnormals).reshape(nt, nn, 3)
            dt_a = z if dtangents is None else _f32(dtangents).reshape(nt, nn, 3)
            dn_e = np.ascontiguousarray(dn_a[:, ni].reshape(-1))
            dt_e = np.ascontiguousarray(dt_a[:, ni].reshape(-1))
            keep += [w, dn_e, dt_e]
            dn, dt = dn_e.ctypes.data_as(C.c_void_p), dt_e.ctypes.data_as(C.c_void_p)
        s = Shading(exp(normals, ni, 3), exp(tangents, ni, 4) if normals is not None else None,
                    exp(uvs, ui, 2), nt, dn, dt, None if w is None else w.ctypes.data_as(C.c_void_p))
        rlib().orc_scene_set_shading(self._h, C.byref(s))

    def set_textures(self, textures):
        """A list of (h, w, 4) float texel arrays (Texture4f)."""
        if not textures:
            rlib().orc_scene_set_textures(self._h, None, None, 0)
            return
        ts = [_f32(t).reshape(t.shape[0], t.shape[1], 4) for t in textures]
        self._texels = np.ascontiguousarray(np.concatenate([t.reshape(-1) for t in ts]))
        sizes = np.array([(t.shape[1], t.shape[0]) for t in ts], np.uint32).reshape(-1)
        rlib().orc_scene_set_textures(self._h, self._texels.ctypes.data_as(C.c_void_p),
                                      sizes.ctypes.data_as(C.c_void_p), len(ts))

    def set_environment(self, rgb):
        if rgb is None:
            rlib().orc_scene_set_environment(self._h, None, 0, 0)
            return
        e = _f32(rgb)
        rlib().orc_scene_set_environment(self._h, e.ctypes.data_as(C.c_void_p), e.shape[1], e.shape[0])

    def shading_attributes(self, prims, b1, b2):
        """Mesh::getIntersectionAttributes per query: ns, ss, ts (n, 3), uv (n, 2), base colour."""
        n = len(prims)
        out = [np.empty((n, 3), np.float32) for _ in range(3)] + [np.empty((n, 2), np.float32),
                                                                  np.empty((n, 3), np.float32)]
        L = rlib()
        for i in range(n):
            r = [np.empty(3, np.float32), np.empty(3, np.float32), np.empty(3, np.float32),
                 np.empty(2, np.float32), np.empty(3, np.float32)]
            L.orc_shading_attributes(self._h, int(prims[i]), float(b1[i]), float(b2[i]), *r)
            for k in range(5):
                out[k][i] = r[k]
        return out

    def environment_sample(self, dirs):
        d = _f32(dirs).reshape(-1, 3)
        out = np.empty_like(d)
        L = rlib()
        for i in range(len(d)):
            o = np.empty(3, np.float32)
            L.orc_environment_sample(self._h, np.ascontiguousarray(d[i]), o)
            out[i] = o
        return out

    def camera_ray(self, x, y):
        o = np.empty(3, np.float32)
        d = np.empty(3, np.float32)
        rlib().orc_camera_ray(self._h, x, y, o, d)
        return o, d

    def light_sampler_sample(self, u, hit, ns):
        p = C.c_float()
        i = rlib().orc_light_sampler_sample(self._h, u, _f32(hit), _f32(ns), C.byref(p))
        return i, p.value

    def light_sampler_pdf(self, light, hit, ns):
        return rlib().orc_light_sampler_pdf(self._h, light, _f32(hit), _f32(ns))

    def light_sample(self, light, u2, hit):
        le, wi = np.empty(3, np.float32), np.empty(3, np.float32)
        d, p = C.c_float(), C.c_float()
        ok = rlib().orc_light_sample(self._h, light, _f32(u2), _f32(hit), le, wi, C.byref(d),
                                     C.byref(p))
        return bool(ok), le, wi, d.value, p.value

    def light_pdf(self, light, hit, wi):
        d = C.c_float()
        p = rlib().orc_light_pdf(self._h, light, _f32(hit), _f32(wi), C.byref(d))
        return p, d.value


def crop_states(width, crop, seed_offset=0):
    """Sampler states of the pixels of window crop = (x0, y0, w, h) of a frame `width` wide,
    each seeded with its frame pixel index (alpine.cpp:259-265)."""
    x0, y0, w, h = crop
    L = rlib()
    st = np.empty((w * h, 4), np.uint32)
    for r in range(h):
        for c in range(w):
            L.orc_sampler_reset(st[r * w + c], ((y0 + r) * width + x0 + c + seed_offset) & 0xFFFFFFFF)
    return st


def sampler_states(n_pixels, seed_offset=0):
    """Per-pixel sampler states seeded with the pixel index (alpine.cpp:259-265)."""
    L = rlib()
    st = np.empty((n_pixels, 4), np.uint32)
    for i in range(n_pixels):
        L.orc_sampler_reset(st[i], (i + seed_offset) & 0xFFFFFFFF)
    return st


def postprocess(accum, total_spp, tonemap=False):
    """Alpine::postProcess without denoising (alpine.cpp:446-483): accum (..., 3) -> u8."""
    a = _f32(accum)
    out = np.empty(a.shape, np.uint8)
    rlib().orc_postprocess(a.reshape(-1), a.size // 3, int(total_spp), int(bool(tonemap)),
                           out.reshape(-1))
    return out


def sampler_sequence(seed, n):
    L = rlib()
    st = np.empty(4, np.uint32)
    L.orc_sampler_reset(st, seed)
    return np.array([L.orc_sampler_get1d(st) for _ in range(n)], np.float32)


def mat_sample(m, wo, u):
    est, wi = np.empty(3, np.float32), np.empty(3, np.float32)
    p = C.c_float()
    ok = rlib().orc_mat_sample(C.byref(m), _f32(wo), _f32(u), est, wi, C.byref(p))
    return bool(ok), est, wi, p.value


def mat_eval(m, wo, wi):
    f = np.empty(3, np.float32)
    p = C.c_float()
    rlib().orc_mat_eval(C.byref(m), _f32(wo), _f32(wi), f, C.byref(p))
    return f, p.value


_ref_render = None


def ref_render_available():
    return os.path.exists(REF_RENDER_LIB_PATH)


def ref_render_lib():
    global _ref_render
    if _ref_render is None:
        L = C.CDLL(REF_RENDER_LIB_PATH)
        fp, u8, u32, i32 = _f32p, _u8p, C.c_uint32, np.ctypeslib.ndpointer(np.int32, flags="C")
        L.ref_sampler_sequence.argtypes = [u32, u32, fp]
        L.ref_camera_rays.argtypes = [fp, fp, fp, C.c_float, C.c_float, fp, u32, fp, fp]
        L.ref_material_sample.argtypes = [u32, fp, fp, fp, fp, u32, u8, fp, fp, fp]
        L.ref_material_eval.argtypes = [u32, fp, fp, fp, fp, u32, fp, fp]
        L.ref_light_sampler.argtypes = [u32, fp, u32, fp, fp, fp, u32, i32, fp]
        L.ref_light_sample.argtypes = [fp, fp, fp, u32, u8, fp, fp, fp, fp]
        L.ref_light_pdf.argtypes = [fp, fp, fp, u32, fp, fp]
        vp, f = C.c_void_p, C.c_float
        L.ref_mesh_attributes.argtypes = [vp, u32, vp, u32, vp, u32, vp, vp, vp, u32, vp, u32, vp,
                                          vp, vp, u32, vp, vp, f, u32, vp, vp, vp, u32, u32, vp,
                                          u32, u32, vp, vp, u32, vp, vp, vp, vp, vp, vp]
        L.ref_environment_sample.argtypes = [vp, u32, u32, vp, u32, vp]
        L.ref_mat4.argtypes = [vp, vp, vp, u32, vp, vp]
        _ref_render = L
    return _ref_render


ADAPTER_LIB_PATH = os.path.join(HERE, "_ref", "libadapter_check.so")
_adapter = None


def adapter_available():
    return os.path.exists(ADAPTER_LIB_PATH)


def adapter_lib():
    """The drop-in adapter (alpine_amd/adapter) beside the reference's Bvh and light samplers
    (oracle/adapter_check.cpp).  Loads alpine_amd's library first so both share one copy."""
    global _adapter
    if _adapter is None:
        import alpine_amd
        alpine_amd.lib()
        L = C.CDLL(ADAPTER_LIB_PATH)
        vp, u32 = C.c_void_p, C.c_uint32
        L.adapter_check_trace.restype = C.c_int
        L.adapter_check_trace.argtypes = [vp, u32, vp, vp, u32, vp, vp, u32, u32] + [vp] * 12
        L.adapter_check_light_sampler.restype = C.c_int
        L.adapter_check_light_sampler.argtypes = [u32, vp, u32, vp, vp, vp, u32, vp, vp, vp, vp]
        L.adapter_check_no_device.restype = C.c_int
        L.adapter_check_no_device.argtypes = []
        _adapter = L
    return _adapter
