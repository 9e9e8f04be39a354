"""alpine_amd -- MI355X (gfx950) back end for alpine's Accelerator / LightSampler plug-ins.

The package is a thin host layer over ``libapg.so`` (hand-written HIP for
gfx950, C ABI in ``include/apg.h``).  It never computes results on the CPU:
when the library or the GPU is missing, the entry points raise.

``GpuAccelerator`` mirrors the reference's ``Accelerator`` interface
(src/accelerators/accelerator.h:13-39): ``append_mesh``, ``get_vertex_buffer``,
``update_shape``, ``update_scene``, ``intersect`` and ``intersect_any`` -- the
last two batched over device-resident ray arrays.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libapg.so")

APG_OK = 0
APG_INVALID_ID = 0xFFFFFFFF
APG_TRACE_STACK_MAX = 64
BUILDER_GPU = 0
BUILDER_HOST = 1

_ERRORS = {
    -1: "invalid argument", -2: "no GPU device", -3: "HIP runtime error",
    -4: "call out of order", -5: "capacity exceeded", -6: "traversal stack overflow",
}


class ApgError(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__(f"apg error {code} ({_ERRORS.get(code, '?')}): {msg}")
        self.code = code


class _Rays(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("ox", "oy", "oz", "dx", "dy", "dz", "tmax")]


class _Hits(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("prim", "t", "b0", "b1")]


class _BuildParams(C.Structure):
    _fields_ = [("leaf_threshold", C.c_uint32), ("builder", C.c_uint32), ("roots", C.c_uint32),
                ("flags", C.c_uint32), ("reserved", C.c_uint32 * 4)]


class FrameTransform(C.Structure):
    _fields_ = [("shape_id", C.c_uint32), ("first", C.c_uint32), ("count", C.c_uint32),
                ("m", C.c_float * 12)]


ROOTS_SCENE, ROOTS_PER_SHAPE = 0, 1
BUILD_NO_STATS = 1


class _SceneStats(C.Structure):
    _fields_ = [
        ("triangles", C.c_uint32), ("shapes", C.c_uint32), ("binary_nodes", C.c_uint32),
        ("binary_leaves", C.c_uint32), ("wide_nodes", C.c_uint32),
        ("max_depth_binary", C.c_uint32), ("sah_cost", C.c_double),
        ("sah_cost_leaf", C.c_double), ("build_ms", C.c_double), ("device_bytes", C.c_uint64),
        ("wide_depth", C.c_uint32), ("reserved", C.c_uint32),
    ]


class Material(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("base_color", C.c_float * 3), ("alpha", C.c_float * 2),
                ("base_color_tex", C.c_uint32), ("normal_tex", C.c_uint32)]


class MeshShading(C.Structure):
    _fields_ = [("normals", C.c_void_p), ("nn", C.c_uint32), ("normal_idx", C.c_void_p),
                ("tangents", C.c_void_p), ("uvs", C.c_void_p), ("nuv", C.c_uint32),
                ("uv_idx", C.c_void_p)]


class Light(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("p0", C.c_float * 3), ("p1", C.c_float * 3),
                ("p2", C.c_float * 3), ("normal", C.c_float * 3), ("color", C.c_float * 3),
                ("power", C.c_float), ("radius", C.c_float)]


class Camera(C.Structure):
    _fields_ = [("eye", C.c_float * 3), ("target", C.c_float * 3), ("up", C.c_float * 3),
                ("fovy", C.c_float), ("aspect", C.c_float)]


class RenderParams(C.Structure):
    _fields_ = [("spp", C.c_uint32), ("max_depth", C.c_uint32), ("rr_depth", C.c_uint32),
                ("fp64_accum", C.c_uint32), ("background", C.c_float * 3), ("guide", C.c_uint32),
                ("train_steps", C.c_uint32), ("guide_warmup", C.c_uint32),
                ("guide_frac", C.c_float), ("graph", C.c_uint32), ("train_samples", C.c_uint32),
                ("guide_depth", C.c_uint32), ("aovs", C.c_uint32), ("count_fetches", C.c_uint32)]


MAT_DIELECTRIC, MAT_CONDUCTOR = 0, 1
LIGHT_DISK, LIGHT_POINT, LIGHT_QUAD, LIGHT_TRIANGLE, LIGHT_POINT_INTENSITY = 0, 1, 2, 3, 4
LS_UNIFORM, LS_POWER, LS_BVH = 0, 1, 2


def _f3(v):
    return (C.c_float * 3)(*[float(x) for x in v])


def material(kind, base_color, alpha, base_color_tex=None, normal_tex=None):
    """A material; base_color_tex / normal_tex name a texture index of set_textures (None: none)."""
    return Material(kind, _f3(base_color), (C.c_float * 2)(*[float(a) for a in alpha]),
                    0 if base_color_tex is None else int(base_color_tex) + 1,
                    0 if normal_tex is None else int(normal_tex) + 1)


def light(kind, p0=(0, 0, 0), p1=(0, 0, 0), p2=(0, 0, 0), normal=(0, 1, 0), color=(1, 1, 1),
          power=1.0, radius=1.0):
    return Light(kind, _f3(p0), _f3(p1), _f3(p2), _f3(normal), _f3(color), float(power),
                 float(radius))


def camera(eye, target, up, fovy, aspect):
    return Camera(_f3(eye), _f3(target), _f3(up), float(fovy), float(aspect))


class TraceConfig(C.Structure):
    _fields_ = [("block_size", C.c_uint32), ("blocks_per_cu", C.c_uint32),
                ("fetch_threshold", C.c_uint32), ("stack_lds", C.c_uint32),
                ("tri_batch", C.c_uint32), ("triangle_test", C.c_uint32), ("stack_limit", C.c_uint32),
                ("reserved", C.c_uint32)]


TRI_REFERENCE, TRI_WATERTIGHT = 0, 1


class _RenderLaunch(C.Structure):
    _fields_ = [("shade_blocks_per_cu", C.c_uint32), ("shadow_blocks_per_cu", C.c_uint32),
                ("resolve_blocks_per_cu", C.c_uint32), ("reserved", C.c_uint32)]


_lib = None

# Every symbol the C ABI headers declare (checked by tests/test_abi.py).
ABI_SYMBOLS = [
    "apg_create", "apg_destroy", "apg_last_error", "apg_mesh_append", "apg_mesh_vertices",
    "apg_mesh_update", "apg_mesh_prim_offset", "apg_mesh_transform", "apg_sphere_append", "apg_scene_build",
    "apg_scene_get_stats", "apg_trace_config_default", "apg_set_trace_config", "apg_intersect",
    "apg_occluded", "apg_intersect_roots", "apg_occluded_roots", "apg_intersect_counted", "apg_scene_triangles", "apg_scene_leaf_keys",
    "apg_scene_wide_nodes", "apg_materials_set", "apg_lights_set", "apg_light_sampler_sample",
    "apg_light_sampler_pdf", "apg_render_reset", "apg_render_reset_slots", "apg_render_set_camera", "apg_render_set_tiles", "apg_render",
    "apg_render_read", "apg_render_postprocess", "apg_render_counters", "apg_render_fetch_counts",
    "apg_guide_param_count",
    "apg_guide_init",
    "apg_guide_params", "apg_guide_forward", "apg_guide_infer", "apg_guide_train_step",
    "apg_guide_train_queries", "apg_render_set_launch", "apg_render_get_launch",
    "apg_render_guide_stats", "apg_mesh_set_shading", "apg_mesh_set_targets", "apg_mesh_morph",
    "apg_textures_set", "apg_render_read_aovs", "apg_render_set_environment",
    "apg_intersect_host", "apg_occluded_host", "apg_light_sampler_set",
    "apg_light_sampler_sample_host", "apg_trace_status", "apg_intersect_noalloc",
    "apg_frame_render", "apg_frame_status",
]
DEBUG_SYMBOLS = ["apg_debug_host_build", "apg_debug_scene_info", "apg_debug_scene_copy",
                 "apg_debug_host_trace", "apg_debug_scene_free", "apg_debug_material",
                 "apg_debug_light_sampler", "apg_debug_light", "apg_debug_shading",
                 "apg_debug_environment", "apg_debug_triangles", "apg_debug_host_trace_limited",
                 "apg_debug_scene_depth", "apg_debug_fail_allocations", "apg_debug_stream_read",
                 "apg_debug_host_rate", "apg_debug_frame_level_cut", "apg_debug_frame_counts"]


def lib():
    """Load libapg.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ApgError(-2, f"{LIB_PATH} is missing: run `python -m alpine_amd.build`")
    # torch ships its own HIP runtime under the same soname (libamdhip64.so.7) as the
    # ROCm one libapg.so links; load torch first so the process has a single runtime
    # (the one torch was built against) whichever module the caller touches first.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    vp, u32, i32 = C.c_void_p, C.c_uint32, C.c_int
    sig = {
        "apg_create": (i32, [C.POINTER(vp), i32]),
        "apg_destroy": (i32, [vp]),
        "apg_last_error": (C.c_char_p, [vp]),
        "apg_mesh_append": (i32, [vp, vp, u32, vp, u32, C.POINTER(u32)]),
        "apg_mesh_vertices": (i32, [vp, u32, C.POINTER(C.POINTER(C.c_float)), C.POINTER(u32)]),
        "apg_mesh_update": (i32, [vp, u32]),
        "apg_mesh_prim_offset": (i32, [vp, u32, C.POINTER(u32)]),
        "apg_mesh_transform": (i32, [vp, u32, vp, u32, u32, vp]),
        "apg_sphere_append": (i32, [vp, vp, u32]),
        "apg_scene_build": (i32, [vp, C.POINTER(_BuildParams)]),
        "apg_scene_get_stats": (i32, [vp, C.POINTER(_SceneStats)]),
        "apg_trace_config_default": (i32, [vp, C.POINTER(TraceConfig)]),
        "apg_set_trace_config": (i32, [vp, C.POINTER(TraceConfig)]),
        "apg_intersect": (i32, [vp, C.POINTER(_Rays), u32, C.POINTER(_Hits), vp]),
        "apg_occluded": (i32, [vp, C.POINTER(_Rays), u32, vp, vp]),
        "apg_intersect_roots": (i32, [vp, C.POINTER(_Rays), vp, u32, C.POINTER(_Hits), vp]),
        "apg_occluded_roots": (i32, [vp, C.POINTER(_Rays), vp, u32, vp, vp]),
        "apg_trace_status": (i32, [vp]),
        "apg_intersect_noalloc": (i32, [vp, C.POINTER(_Rays), u32, C.POINTER(_Hits), vp]),
        "apg_intersect_counted": (i32, [vp, C.POINTER(_Rays), u32, C.POINTER(_Hits), i32, vp,
                                        C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), vp]),
        "apg_scene_triangles": (i32, [vp, vp, vp, u32]),
        "apg_scene_leaf_keys": (i32, [vp, vp, u32]),
        "apg_scene_wide_nodes": (i32, [vp, vp, u32]),
        "apg_materials_set": (i32, [vp, vp, u32, vp, u32]),
        "apg_lights_set": (i32, [vp, vp, u32, u32]),
        "apg_light_sampler_sample": (i32, [vp, vp, vp, vp, u32, vp, vp, vp]),
        "apg_light_sampler_pdf": (i32, [vp, vp, vp, vp, u32, vp, vp]),
        "apg_light_sampler_sample_host": (i32, [vp, vp, vp, vp, u32, vp, vp]),
        "apg_intersect_host": (i32, [vp, vp, u32, vp, vp, vp, vp]),
        "apg_occluded_host": (i32, [vp, vp, vp, u32, vp]),
        "apg_render_reset": (i32, [vp, u32, u32, u32]),
        "apg_render_reset_slots": (i32, [vp, u32, u32, u32, u32]),
        "apg_render_set_camera": (i32, [vp, C.POINTER(Camera)]),
        "apg_render_set_tiles": (i32, [vp, vp, vp, u32, u32, u32]),
        "apg_render": (i32, [vp, C.POINTER(RenderParams), vp]),
        "apg_frame_render": (i32, [vp, vp, u32, C.POINTER(_BuildParams), C.POINTER(RenderParams), u32, vp]),
        "apg_frame_status": (i32, [vp]),
        "apg_debug_frame_level_cut": (i32, [vp, u32]),
        "apg_debug_frame_counts": (i32, [vp, C.POINTER(u32), C.POINTER(u32), C.POINTER(u32)]),
        "apg_render_read": (i32, [vp, vp, vp, C.POINTER(u32)]),
        "apg_render_read_aovs": (i32, [vp, vp, vp]),
        "apg_render_set_environment": (i32, [vp, vp, u32, u32]),
        "apg_textures_set": (i32, [vp, vp, vp, u32]),
        "apg_mesh_set_shading": (i32, [vp, u32, C.POINTER(MeshShading)]),
        "apg_mesh_set_targets": (i32, [vp, u32, u32, vp, vp, vp]),
        "apg_mesh_morph": (i32, [vp, u32, vp, vp]),
        "apg_render_postprocess": (i32, [vp, i32, vp, vp]),
        "apg_render_fetch_counts": (i32, [vp] + [C.POINTER(C.c_uint64)] * 4),
        "apg_render_counters": (i32, [vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                      C.POINTER(C.c_uint64)]),
        "apg_guide_param_count": (i32, []),
        "apg_render_guide_stats": (i32, [vp, C.POINTER(u32), C.POINTER(C.c_float)]),
        "apg_guide_init": (i32, [vp, u32, C.c_float]),
        "apg_guide_params": (i32, [vp, vp, i32]),
        "apg_guide_forward": (i32, [vp, vp, u32, vp, vp]),
        "apg_guide_infer": (i32, [vp, vp, u32, vp, vp]),
        "apg_guide_train_step": (i32, [vp, vp, vp, vp, u32, i32, C.POINTER(C.c_float), vp, vp]),
        "apg_guide_train_queries": (i32, [vp, vp, vp, vp, u32, i32, C.POINTER(C.c_float), vp, vp]),
        "apg_render_set_launch": (i32, [vp, vp]),
        "apg_render_get_launch": (i32, [vp, vp]),
        "apg_debug_host_build": (i32, [vp, u32, vp, u32, u32, C.POINTER(vp)]),
        "apg_debug_scene_info": (i32, [vp, C.POINTER(u32), C.POINTER(u32), C.POINTER(C.c_double),
                                       C.POINTER(C.c_double), C.POINTER(u32), C.POINTER(u32)]),
        "apg_debug_scene_copy": (i32, [vp, vp, vp, vp, vp]),
        "apg_debug_host_trace": (i32, [vp, vp, vp, vp, u32, i32, vp, vp, vp, vp, vp,
                                       C.POINTER(u32)]),
        "apg_debug_scene_free": (None, [vp]),
        "apg_debug_host_trace_limited": (i32, [vp, vp, vp, vp, u32, i32, u32, vp, vp, vp, vp, vp,
                                               C.POINTER(u32)]),
        "apg_debug_scene_depth": (i32, [vp, C.POINTER(u32)]),
        "apg_debug_fail_allocations": (i32, [vp, i32]),
        "apg_debug_stream_read": (i32, [vp, C.c_uint64, vp, u32, vp]),
        "apg_debug_host_rate": (i32, [vp, vp, u32, u32, u32, i32, vp, vp, vp, C.POINTER(C.c_double)]),
        "apg_debug_material": (i32, [C.POINTER(Material)] + [vp] * 3 + [u32] + [vp] * 6),
        "apg_debug_light_sampler": (i32, [vp, u32, u32, vp, vp, vp, u32, vp, vp, vp]),
        "apg_debug_light": (i32, [C.POINTER(Light), vp, vp, vp, u32] + [vp] * 7),
        "apg_debug_shading": (i32, [vp, u32, u32, C.POINTER(MeshShading), u32, vp, vp, vp,
                                    C.POINTER(Material), vp, vp, u32, vp, vp, vp, u32] + [vp] * 5),
        "apg_debug_environment": (i32, [vp, u32, u32, vp, u32, vp]),
        "apg_debug_triangles": (i32, [vp, u32, vp, vp, vp, u32, i32, vp, vp, vp, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if isinstance(a, np.ndarray) else C.c_void_p(a.data_ptr())


def _check_cuda(t, name):
    import torch
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise ApgError(-1, f"{name} must be a CUDA (HIP) tensor")
    if not t.is_contiguous():
        raise ApgError(-1, f"{name} must be contiguous")


class GpuAccelerator:
    """The ``gpu`` Accelerator (reference interface: accelerator.h:13-39)."""

    INVALID_SHAPE_ID = APG_INVALID_ID

    def __init__(self, device: int = 0):
        self._lib = lib()
        h = C.c_void_p()
        rc = self._lib.apg_create(C.byref(h), device)
        if rc != APG_OK:
            raise ApgError(rc, self._lib.apg_last_error(None).decode())
        self._h = h
        self.device = device
        self._keep = []

    def close(self):
        if getattr(self, "_h", None):
            self._lib.apg_destroy(self._h)
            self._h = None

    __del__ = close

    def _rc(self, rc):
        if rc != APG_OK:
            raise ApgError(rc, self._lib.apg_last_error(self._h).decode())

    # ---- geometry ----
    def append_mesh(self, vertices, prims) -> int:
        v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
        t = np.ascontiguousarray(prims, np.uint32).reshape(-1, 3)
        sid = C.c_uint32()
        self._rc(self._lib.apg_mesh_append(self._h, _ptr(v), len(v), _ptr(t), len(t),
                                           C.byref(sid)))
        return sid.value

    def append_sphere(self, vertices):
        v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 4)
        self._rc(self._lib.apg_sphere_append(self._h, _ptr(v), len(v)))

    def get_vertex_buffer(self, shape_id: int) -> np.ndarray:
        p = C.POINTER(C.c_float)()
        nv = C.c_uint32()
        self._rc(self._lib.apg_mesh_vertices(self._h, shape_id, C.byref(p), C.byref(nv)))
        return np.ctypeslib.as_array(p, shape=(nv.value, 3))

    def update_shape(self, shape_id: int):
        self._rc(self._lib.apg_mesh_update(self._h, shape_id))

    def transform_vertices(self, shape_id, m34, first=0, count=None, stream=None):
        """Rigid 3 x 4 transform of a vertex range, applied on the GPU (no upload)."""
        m = np.ascontiguousarray(m34, np.float32).reshape(12)
        if count is None:
            count = self.vertex_count(shape_id) - first
        self._rc(self._lib.apg_mesh_transform(self._h, shape_id, _ptr(m), first, count,
                                              self._stream(stream)))

    def set_shading(self, shape_id, normals=None, normal_idx=None, tangents=None, uvs=None,
                    uv_idx=None):
        """Shading attributes of a mesh (Mesh::Data normals / tangents / uvs and their index
        buffers, shapes/mesh.h:16-33); None leaves an attribute out."""
        keep = []

        def arr(a, dtype, cols):
            if a is None:
                return None, 0
            a = np.ascontiguousarray(a, dtype).reshape(-1, cols)
            keep.append(a)
            return _ptr(a), len(a)

        n, nn = arr(normals, np.float32, 3)
        ni, _ = arr(normal_idx, np.uint32, 3)
        t, nt = arr(tangents, np.float32, 4)
        if tangents is not None and nt != nn:
            raise ApgError(-1, "tangents must have one row per normal")
        u, nuv = arr(uvs, np.float32, 2)
        ui, _ = arr(uv_idx, np.uint32, 3)
        s = MeshShading(n, nn, ni, t, u, nuv, ui)
        self._rc(self._lib.apg_mesh_set_shading(self._h, shape_id, C.byref(s)))

    def set_targets(self, shape_id, dpos, dnormals=None, dtangents=None):
        """Morph targets (Mesh::Data::Target): dpos (ntargets, nv, 3), optional normal and
        tangent deltas (ntargets, nn, 3).  The current vertices become the blend base."""
        d = np.ascontiguousarray(dpos, np.float32)
        nt = 0 if d.size == 0 else d.shape[0]
        dn = None if dnormals is None else np.ascontiguousarray(dnormals, np.float32)
        dt = None if dtangents is None else np.ascontiguousarray(dtangents, np.float32)
        self._rc(self._lib.apg_mesh_set_targets(self._h, shape_id, nt, _ptr(d) if nt else None,
                                                None if dn is None else _ptr(dn),
                                                None if dt is None else _ptr(dt)))

    def morph(self, shape_id, weights, stream=None):
        """Mesh::update on the GPU: vertices = base + sum_j dpos_j * weights[j]."""
        w = np.ascontiguousarray(weights, np.float32).reshape(-1)
        self._rc(self._lib.apg_mesh_morph(self._h, shape_id, _ptr(w), self._stream(stream)))

    def vertex_count(self, shape_id):
        p = C.POINTER(C.c_float)()
        nv = C.c_uint32()
        self._rc(self._lib.apg_mesh_vertices(self._h, shape_id, C.byref(p), C.byref(nv)))
        return nv.value

    def prim_offset(self, shape_id: int) -> int:
        o = C.c_uint32()
        self._rc(self._lib.apg_mesh_prim_offset(self._h, shape_id, C.byref(o)))
        return o.value

    def update_scene(self, leaf_threshold: int = 1, builder: int = BUILDER_GPU,
                     per_shape: bool = False, stats: bool = True):
        """Rebuild (Accelerator::updateScene).  per_shape: an independent tree per shape
        (GPU builder), wide node s the root of shape s; stats=False skips the host-side
        tree statistics (per-frame rebuilds)."""
        p = _BuildParams(leaf_threshold, builder, ROOTS_PER_SHAPE if per_shape else ROOTS_SCENE,
                         0 if stats else BUILD_NO_STATS)
        self._rc(self._lib.apg_scene_build(self._h, C.byref(p)))

    def stats(self) -> dict:
        s = _SceneStats()
        self._rc(self._lib.apg_scene_get_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in _SceneStats._fields_}

    def trace_config(self) -> TraceConfig:
        """The launch shape in use (the library default until set_trace_config)."""
        cur = getattr(self, "_cfg", None)
        if cur is not None:
            return TraceConfig.from_buffer_copy(cur)
        c = TraceConfig()
        self._rc(self._lib.apg_trace_config_default(self._h, C.byref(c)))
        return c

    def set_trace_config(self, cfg: TraceConfig):
        self._rc(self._lib.apg_set_trace_config(self._h, C.byref(cfg)))
        self._cfg = TraceConfig.from_buffer_copy(cfg)

    def set_triangle_test(self, kind: int):
        """TRI_REFERENCE (the reference's Moller-Trumbore, bit-exact hits) or TRI_WATERTIGHT."""
        c = self.trace_config()
        c.triangle_test = kind
        self.set_trace_config(c)

    # ---- traversal ----
    @staticmethod
    def _rays(org, dir, tmax=None):
        for name, t in (("org", org), ("dir", dir)):
            _check_cuda(t, name)
        if org.shape[0] != 3 or dir.shape[0] != 3:
            raise ApgError(-1, "rays are SoA: org and dir must have shape (3, n)")
        r = _Rays(org[0].data_ptr(), org[1].data_ptr(), org[2].data_ptr(),
                  dir[0].data_ptr(), dir[1].data_ptr(), dir[2].data_ptr(),
                  tmax.data_ptr() if tmax is not None else None)
        return r

    @staticmethod
    def _stream(stream):
        import torch
        s = stream if stream is not None else torch.cuda.current_stream()
        return C.c_void_p(s.cuda_stream)

    def intersect(self, org, dir, out=None, stream=None):
        """Closest hit for rays org/dir of shape (3, n) float32 on the GPU.

        Returns (prim uint32-as-int32, t, b0, b1) device tensors.
        """
        import torch
        n = org.shape[1]
        rays = self._rays(org, dir)
        if out is None:
            out = (torch.empty(n, dtype=torch.int32, device=org.device),
                   torch.empty(n, dtype=torch.float32, device=org.device),
                   torch.empty(n, dtype=torch.float32, device=org.device),
                   torch.empty(n, dtype=torch.float32, device=org.device))
        hits = _Hits(*(t.data_ptr() for t in out))
        self._rc(self._lib.apg_intersect(self._h, C.byref(rays), n, C.byref(hits),
                                         self._stream(stream)))
        return out

    def intersect_roots(self, org, dir, roots, stream=None):
        """Closest hit of ray i in the tree of shape roots[i] (per-shape forest)."""
        import torch
        _check_cuda(roots, "roots")
        n = org.shape[1]
        rays = self._rays(org, dir)
        out = (torch.empty(n, dtype=torch.int32, device=org.device),
               torch.empty(n, dtype=torch.float32, device=org.device),
               torch.empty(n, dtype=torch.float32, device=org.device),
               torch.empty(n, dtype=torch.float32, device=org.device))
        hits = _Hits(*(t.data_ptr() for t in out))
        self._rc(self._lib.apg_intersect_roots(self._h, C.byref(rays), C.c_void_p(roots.data_ptr()), n,
                                               C.byref(hits), self._stream(stream)))
        return out

    def intersect_any_roots(self, org, dir, roots, tmax=None, stream=None):
        import torch
        _check_cuda(roots, "roots")
        n = org.shape[1]
        rays = self._rays(org, dir, tmax)
        out = torch.empty(n, dtype=torch.uint8, device=org.device)
        self._rc(self._lib.apg_occluded_roots(self._h, C.byref(rays), C.c_void_p(roots.data_ptr()), n,
                                              C.c_void_p(out.data_ptr()), self._stream(stream)))
        return out

    def intersect_any(self, org, dir, tmax=None, out=None, stream=None):
        """Any hit with t < tmax (FLT_MAX when None); returns a uint8 device tensor."""
        import torch
        n = org.shape[1]
        if tmax is not None:
            _check_cuda(tmax, "tmax")
        rays = self._rays(org, dir, tmax)
        if out is None:
            out = torch.empty(n, dtype=torch.uint8, device=org.device)
        self._rc(self._lib.apg_occluded(self._h, C.byref(rays), n, C.c_void_p(out.data_ptr()),
                                        self._stream(stream)))
        return out

    def trace_status(self):
        """Wait for the traversals issued so far; raise ApgError(-6) if one overflowed its stack."""
        self._rc(self._lib.apg_trace_status(self._h))

    def intersect_counted(self, org, dir, tmax=None, any_hit=False, stream=None):
        """Instrumented run: results plus total node and triangle fetches."""
        import torch
        n = org.shape[1]
        rays = self._rays(org, dir, tmax)
        nodes, tris = C.c_uint64(), C.c_uint64()
        if any_hit:
            occ = torch.empty(n, dtype=torch.uint8, device=org.device)
            self._rc(self._lib.apg_intersect_counted(self._h, C.byref(rays), n, None, 1,
                                                     C.c_void_p(occ.data_ptr()), C.byref(nodes),
                                                     C.byref(tris), self._stream(stream)))
            return occ, nodes.value, tris.value
        out = (torch.empty(n, dtype=torch.int32, device=org.device),
               torch.empty(n, dtype=torch.float32, device=org.device),
               torch.empty(n, dtype=torch.float32, device=org.device),
               torch.empty(n, dtype=torch.float32, device=org.device))
        hits = _Hits(*(t.data_ptr() for t in out))
        self._rc(self._lib.apg_intersect_counted(self._h, C.byref(rays), n, C.byref(hits), 0,
                                                 None, C.byref(nodes), C.byref(tris),
                                                 self._stream(stream)))
        return out, nodes.value, tris.value

    # ---- shading data and the path tracer ----
    def set_materials(self, materials, tri_material):
        arr = (Material * len(materials))(*materials)
        tm = np.ascontiguousarray(tri_material, np.uint32)
        self._keep_mats = arr
        self._rc(self._lib.apg_materials_set(self._h, C.cast(arr, C.c_void_p), len(materials),
                                             _ptr(tm), len(tm)))

    def set_lights(self, lights, sampler=LS_BVH):
        arr = (Light * max(1, len(lights)))(*lights)
        self._rc(self._lib.apg_lights_set(self._h, C.cast(arr, C.c_void_p), len(lights), sampler))

    def light_sampler_sample(self, u, hit, ns, stream=None):
        """LightSampler::sample for device tensors u (n), hit/ns (n, 3); returns (light, pdf)."""
        import torch
        n = u.shape[0]
        light = torch.empty(n, dtype=torch.int32, device=u.device)
        pdf = torch.empty(n, dtype=torch.float32, device=u.device)
        self._rc(self._lib.apg_light_sampler_sample(self._h, _ptr(u), _ptr(hit), _ptr(ns), n,
                                                    _ptr(light), _ptr(pdf), self._stream(stream)))
        return light, pdf

    def light_sampler_sample_host(self, u, hit, ns):
        """LightSampler::sample for host arrays u (n), hit/ns (n, 3), synchronous; returns
        numpy (light, pdf)."""
        u = np.ascontiguousarray(u, np.float32)
        hit = np.ascontiguousarray(hit, np.float32)
        ns = np.ascontiguousarray(ns, np.float32)
        n = u.shape[0]
        assert hit.shape == (n, 3) and ns.shape == (n, 3)
        light = np.empty(n, np.int32)
        pdf = np.empty(n, np.float32)
        self._rc(self._lib.apg_light_sampler_sample_host(self._h, _ptr(u), _ptr(hit), _ptr(ns), n,
                                                         _ptr(light), _ptr(pdf)))
        return light, pdf

    def light_sampler_pdf(self, light, hit, ns, stream=None):
        import torch
        n = light.shape[0]
        pdf = torch.empty(n, dtype=torch.float32, device=light.device)
        self._rc(self._lib.apg_light_sampler_pdf(self._h, _ptr(light), _ptr(hit), _ptr(ns), n,
                                                 _ptr(pdf), self._stream(stream)))
        return pdf

    def render_reset(self, width, height, seed=0, slots=1):
        """Clear the accumulation; `slots` samples per pixel are traced side by side per pass."""
        self.width, self.height = width, height
        if slots == 1:
            self._rc(self._lib.apg_render_reset(self._h, width, height, seed))
        else:
            self._rc(self._lib.apg_render_reset_slots(self._h, width, height, seed, slots))

    def set_camera(self, cam):
        self._rc(self._lib.apg_render_set_camera(self._h, C.byref(cam)))

    def set_tiles(self, cams, tile_width, tile_height, roots=None):
        """Tiled frame: tile t (row-major) seen through cams[t], traced in tree roots[t]."""
        arr = (Camera * max(1, len(cams)))(*cams)
        r = None if roots is None else np.ascontiguousarray(roots, np.uint32)
        self._rc(self._lib.apg_render_set_tiles(self._h, C.cast(arr, C.c_void_p),
                                                None if r is None else _ptr(r), len(cams),
                                                tile_width, tile_height))

    def render(self, spp, max_depth, rr_depth=0, background=(0.0, 0.0, 0.0), fp64=False,
               guide=0, train_steps=4, guide_warmup=1, guide_frac=0.5, graph=False,
               train_samples=0, guide_depth=0, aovs=False, count_fetches=False, stream=None):
        p = RenderParams(spp, max_depth, rr_depth, 1 if fp64 else 0, _f3(background), guide,
                         train_steps, guide_warmup, guide_frac, 1 if graph else 0, train_samples,
                         guide_depth, 1 if aovs else 0, 1 if count_fetches else 0)
        self._rc(self._lib.apg_render(self._h, C.byref(p), self._stream(stream)))

    def frame_render(self, transforms, spp, max_depth, rr_depth=0, seed=0, leaf_threshold=1,
                     per_shape=False, stats=False, background=(0.0, 0.0, 0.0), fp64=False,
                     stream=None):
        """One animated frame (apg_frame_render): transforms = [(shape_id, first, count, m34)],
        then the GPU rebuild, a reset at the current size with `seed` and `spp` samples; the
        first frame runs eagerly, later ones replay one captured HIP graph."""
        arr = (FrameTransform * max(1, len(transforms)))()
        for i, (sid, first, count, m34) in enumerate(transforms):
            arr[i].shape_id, arr[i].first, arr[i].count = int(sid), int(first), int(count)
            arr[i].m = (C.c_float * 12)(*[float(x) for x in np.asarray(m34, np.float32).reshape(12)])
        bp = _BuildParams(leaf_threshold, BUILDER_GPU, ROOTS_PER_SHAPE if per_shape else ROOTS_SCENE,
                          0 if stats else BUILD_NO_STATS)
        rp = RenderParams(spp, max_depth, rr_depth, 1 if fp64 else 0, _f3(background), 0, 4, 1, 0.5,
                          0, 0, 0, 0, 0)
        self._rc(self._lib.apg_frame_render(self._h, C.cast(arr, C.c_void_p), len(transforms),
                                            C.byref(bp), C.byref(rp), seed, self._stream(stream)))

    def frame_status(self) -> int:
        """Waits for the last replayed frame: 0, or the APG_E_CAPACITY code when its rebuild ran
        out of levels (that frame shows the previous tree)."""
        return int(self._lib.apg_frame_status(self._h))

    def frame_counts(self):
        """(eager frames, captures, replays) of frame_render (test hook)."""
        e, c, r = C.c_uint32(), C.c_uint32(), C.c_uint32()
        self._rc(self._lib.apg_debug_frame_counts(self._h, C.byref(e), C.byref(c), C.byref(r)))
        return e.value, c.value, r.value

    def frame_level_cut(self, cut):
        """Test hook: captured frames run `cut` binary levels fewer than they need."""
        self._rc(self._lib.apg_debug_frame_level_cut(self._h, int(cut)))

    def render_read_aovs(self):
        """Albedo and normal render-target sums (height, width, 3) of the aovs=True passes."""
        alb = np.empty((self.height, self.width, 3), np.float32)
        nrm = np.empty((self.height, self.width, 3), np.float32)
        self._rc(self._lib.apg_render_read_aovs(self._h, _ptr(alb), _ptr(nrm)))
        return alb, nrm

    def set_textures(self, textures):
        """Textures (Texture4f): a list of (h, w, 4) float arrays of linear texel values."""
        if not textures:
            self._rc(self._lib.apg_textures_set(self._h, None, None, 0))
            return
        ts = [np.ascontiguousarray(t, np.float32).reshape(t.shape[0], t.shape[1], 4) for t in textures]
        texels = np.ascontiguousarray(np.concatenate([t.reshape(-1) for t in ts]))
        sizes = np.array([(t.shape[1], t.shape[0]) for t in ts], np.uint32).reshape(-1)
        self._rc(self._lib.apg_textures_set(self._h, _ptr(texels), _ptr(sizes), len(ts)))

    def set_environment(self, rgb):
        """Environment map on a miss: (h, w, 3) float array, or None for the background."""
        if rgb is None:
            self._rc(self._lib.apg_render_set_environment(self._h, None, 0, 0))
            return
        e = np.ascontiguousarray(rgb, np.float32)
        self._rc(self._lib.apg_render_set_environment(self._h, _ptr(e), e.shape[1], e.shape[0]))

    def render_launch(self):
        """The wavefront stages' workgroups per CU (apg_render_launch)."""
        r = _RenderLaunch()
        self._rc(self._lib.apg_render_get_launch(self._h, C.byref(r)))
        return {"shade": r.shade_blocks_per_cu, "shadow": r.shadow_blocks_per_cu, "resolve": r.resolve_blocks_per_cu}

    def set_render_launch(self, shade=0, shadow=0, resolve=0):
        r = _RenderLaunch(shade, shadow, resolve, 0)
        self._rc(self._lib.apg_render_set_launch(self._h, C.byref(r)))

    def guide_stats(self):
        n, loss = C.c_uint32(), C.c_float()
        self._rc(self._lib.apg_render_guide_stats(self._h, C.byref(n), C.byref(loss)))
        return n.value, loss.value

    def render_read(self, fp64=False):
        rgb = np.empty((self.height, self.width, 3), np.float32)
        r64 = np.empty((self.height, self.width, 3), np.float64) if fp64 else None
        spp = C.c_uint32()
        self._rc(self._lib.apg_render_read(self._h, _ptr(rgb), None if r64 is None else _ptr(r64),
                                           C.byref(spp)))
        return rgb, r64, spp.value

    def postprocess(self, tonemap=False, stream=None):
        """Display frame (height x width x 3 uint8 CUDA tensor): Alpine::postProcess."""
        import torch
        frame = torch.empty((self.height, self.width, 3), dtype=torch.uint8, device="cuda")
        self._rc(self._lib.apg_render_postprocess(self._h, int(bool(tonemap)), frame.data_ptr(),
                                                  self._stream(stream)))
        return frame

    def render_fetch_counts(self):
        """Node and triangle fetches of the count_fetches passes since render_reset."""
        v = [C.c_uint64() for _ in range(4)]
        self._rc(self._lib.apg_render_fetch_counts(self._h, *[C.byref(x) for x in v]))
        return dict(zip(["closest_nodes", "closest_tris", "shadow_nodes", "shadow_tris"], [x.value for x in v]))

    def render_counters(self):
        a, b, c = C.c_uint64(), C.c_uint64(), C.c_uint64()
        self._rc(self._lib.apg_render_counters(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return {"closest_rays": a.value, "shadow_rays": b.value, "vertices": c.value}

    # ---- learned path guiding ----
    def guide_init(self, seed=1, lr=1e-3):
        self._rc(self._lib.apg_guide_init(self._h, seed, lr))

    def guide_params(self, params=None):
        n = self._lib.apg_guide_param_count()
        if params is None:
            out = np.empty(n, np.float32)
            self._rc(self._lib.apg_guide_params(self._h, _ptr(out), 0))
            return out
        p = np.ascontiguousarray(params, np.float32)
        assert p.shape == (n,)
        self._rc(self._lib.apg_guide_params(self._h, _ptr(p), 1))

    def guide_forward(self, feat, stream=None):
        import torch
        n = feat.shape[0]
        out = torch.empty((n, 48), dtype=torch.float32, device=feat.device)
        self._rc(self._lib.apg_guide_forward(self._h, _ptr(feat), n, _ptr(out), self._stream(stream)))
        return out

    def guide_infer(self, q, stream=None):
        """Proposal parameters (n x 48) from n x 9 queries (p01, n, wo), encoding fused in."""
        import torch
        n = q.shape[0]
        out = torch.empty((n, 48), dtype=torch.float32, device=q.device)
        self._rc(self._lib.apg_guide_infer(self._h, _ptr(q), n, _ptr(out), self._stream(stream)))
        return out

    def guide_train_queries(self, q, dirs, weight, apply=True, want_grad=False, want_loss=True, stream=None):
        """One training step from n x 9 queries (encoding fused into the training kernel).
        Returns (loss, grad); with want_loss=False and want_grad=False the step is only
        enqueued on the stream and (None, None) comes back at once."""
        return self._train(self._lib.apg_guide_train_queries, q, dirs, weight, apply, want_grad, want_loss, stream)

    def guide_train_step(self, feat, dirs, weight, apply=True, want_grad=False, want_loss=True, stream=None):
        """One training step from n x 96 features; see guide_train_queries."""
        return self._train(self._lib.apg_guide_train_step, feat, dirs, weight, apply, want_grad, want_loss, stream)

    def _train(self, fn, x, dirs, weight, apply, want_grad, want_loss, stream):
        loss = C.c_float() if want_loss else None
        g = np.empty(self._lib.apg_guide_param_count(), np.float32) if want_grad else None
        self._rc(fn(self._h, _ptr(x), _ptr(dirs), _ptr(weight), x.shape[0], 1 if apply else 0,
                    None if loss is None else C.byref(loss), None if g is None else _ptr(g),
                    self._stream(stream)))
        return (None if loss is None else loss.value), g

    # ---- introspection ----
    def leaf_keys(self) -> np.ndarray:
        n = self.stats()["triangles"]
        out = np.empty(n, np.uint32)
        self._rc(self._lib.apg_scene_leaf_keys(self._h, _ptr(out), n))
        return out

    def triangles(self):
        n = self.stats()["triangles"]
        tri9 = np.empty((n, 9), np.float32)
        ids = np.empty(n, np.uint32)
        self._rc(self._lib.apg_scene_triangles(self._h, _ptr(tri9), _ptr(ids), n))
        return tri9, ids

    def host_rate(self, rays6, nthreads, rays_per_call=1, any_hit=False):
        """The per-ray host entry points from `nthreads` native threads (apg_debug_host_rate):
        (results, seconds); results are (prim, t) or the occlusion bytes."""
        r = np.ascontiguousarray(rays6, np.float32).reshape(-1, 6)
        n = len(r)
        sec = C.c_double()
        if any_hit:
            occ = np.empty(n, np.uint8)
            self._rc(self._lib.apg_debug_host_rate(self._h, _ptr(r), n, nthreads, rays_per_call, 1, None, None,
                                                   _ptr(occ), C.byref(sec)))
            return occ, sec.value
        prim = np.empty(n, np.uint32)
        t = np.empty(n, np.float32)
        self._rc(self._lib.apg_debug_host_rate(self._h, _ptr(r), n, nthreads, rays_per_call, 0, _ptr(prim), _ptr(t),
                                               None, C.byref(sec)))
        return (prim, t), sec.value

    def wide_nodes(self) -> np.ndarray:
        n = self.stats()["wide_nodes"]
        out = np.empty((n, 16), np.uint32)
        self._rc(self._lib.apg_scene_wide_nodes(self._h, _ptr(out), n))
        return out


class DebugScene:
    """Host build + host traversal replay (include/apg_debug.h); CPU tests only."""

    def __init__(self, verts, tris, leaf_threshold=4):
        L = lib()
        v = np.ascontiguousarray(verts, np.float32).reshape(-1, 3)
        t = np.ascontiguousarray(tris, np.uint32).reshape(-1, 3)
        h = C.c_void_p()
        rc = L.apg_debug_host_build(_ptr(v), len(v), _ptr(t), len(t), leaf_threshold,
                                    C.byref(h))
        if rc != APG_OK:
            raise ApgError(rc, "debug host build failed")
        self._h = h
        self.nt = len(t)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().apg_debug_scene_free(self._h)
            self._h = None

    def info(self):
        nn, nt, bn, md = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32()
        sah, sahl = C.c_double(), C.c_double()
        lib().apg_debug_scene_info(self._h, C.byref(nn), C.byref(nt), C.byref(sah),
                                   C.byref(sahl), C.byref(bn), C.byref(md))
        return {"wide_nodes": nn.value, "triangles": nt.value, "sah_cost": sah.value,
                "sah_cost_leaf": sahl.value, "binary_nodes": bn.value, "max_depth": md.value}

    def wide_depth(self):
        d = C.c_uint32()
        lib().apg_debug_scene_depth(self._h, C.byref(d))
        return d.value

    def trace_limited(self, org, dir, stack_limit, any_hit=False):
        """Host trace with the kernel's bounded stack: (results, rays that overflowed)."""
        org = np.ascontiguousarray(org, np.float32)
        dir = np.ascontiguousarray(dir, np.float32)
        n = len(org)
        over = C.c_uint32()
        if any_hit:
            occ = np.empty(n, np.uint8)
            rc = lib().apg_debug_host_trace_limited(self._h, _ptr(org), _ptr(dir), None, n, 1, stack_limit,
                                                    None, None, None, None, _ptr(occ), C.byref(over))
            if rc != APG_OK:
                raise ApgError(rc, "apg_debug_host_trace_limited failed")
            return occ, over.value
        prim = np.empty(n, np.uint32)
        t, b0, b1 = (np.empty(n, np.float32) for _ in range(3))
        rc = lib().apg_debug_host_trace_limited(self._h, _ptr(org), _ptr(dir), None, n, 0, stack_limit, _ptr(prim),
                                                _ptr(t), _ptr(b0), _ptr(b1), None, C.byref(over))
        if rc != APG_OK:
            raise ApgError(rc, "apg_debug_host_trace_limited failed")
        return (prim, t, b0, b1), over.value

    def arrays(self):
        i = self.info()
        nodes = np.empty((i["wide_nodes"], 16), np.uint32)
        tri9 = np.empty((i["triangles"], 9), np.float32)
        ids = np.empty(i["triangles"], np.uint32)
        keys = np.empty(self.nt, np.uint32)
        lib().apg_debug_scene_copy(self._h, _ptr(nodes), _ptr(tri9), _ptr(ids), _ptr(keys))
        return nodes, tri9, ids, keys

    def trace(self, org, dir, tmax=None, any_hit=False):
        org = np.ascontiguousarray(org, np.float32)
        dir = np.ascontiguousarray(dir, np.float32)
        n = len(org)
        tm = None if tmax is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(tmax, np.float32), (n,)))
        depth = C.c_uint32()
        if any_hit:
            occ = np.empty(n, np.uint8)
            lib().apg_debug_host_trace(self._h, _ptr(org), _ptr(dir),
                                       None if tm is None else _ptr(tm), n, 1, None, None, None,
                                       None, _ptr(occ), C.byref(depth))
            return occ, depth.value
        prim = np.empty(n, np.uint32)
        t = np.empty(n, np.float32)
        b0 = np.empty(n, np.float32)
        b1 = np.empty(n, np.float32)
        lib().apg_debug_host_trace(self._h, _ptr(org), _ptr(dir),
                                   None if tm is None else _ptr(tm), n, 0, _ptr(prim), _ptr(t),
                                   _ptr(b0), _ptr(b1), None, C.byref(depth))
        return (prim, t, b0, b1), depth.value


def debug_shading(idx, nv, queries_prim, queries_b, ng, mat, normals=None, normal_idx=None,
                  tangents=None, uvs=None, uv_idx=None, dnormals=None, dtangents=None,
                  weights=None, textures=()):
    """Host build of the device shading-attribute steps (test hook, apg_debug.h): per query
    (prim, (b1, b2), ng) the shading normal, its basis, uv and the base colour."""
    keep = []

    def arr(a, dtype, cols):
        if a is None:
            return None, 0
        a = np.ascontiguousarray(a, dtype).reshape(-1, cols)
        keep.append(a)
        return _ptr(a), len(a)

    n, nn = arr(normals, np.float32, 3)
    ni, _ = arr(normal_idx, np.uint32, 3)
    t, _ = arr(tangents, np.float32, 4)
    u, nuv = arr(uvs, np.float32, 2)
    ui, _ = arr(uv_idx, np.uint32, 3)
    sh = MeshShading(n, nn, ni, t, u, nuv, ui)
    tri = np.ascontiguousarray(idx, np.uint32).reshape(-1, 3)
    w = None if weights is None else np.ascontiguousarray(weights, np.float32).reshape(-1)
    nt_ = 0 if w is None else len(w)
    dn = None if dnormals is None else np.ascontiguousarray(dnormals, np.float32)
    dt = None if dtangents is None else np.ascontiguousarray(dtangents, np.float32)
    ts = [np.ascontiguousarray(x, np.float32) for x in textures]
    texels = np.ascontiguousarray(np.concatenate([x.reshape(-1) for x in ts])) if ts else np.zeros(4, np.float32)
    sizes = np.array([(x.shape[1], x.shape[0]) for x in ts], np.uint32).reshape(-1) if ts else np.zeros(2, np.uint32)
    qp = np.ascontiguousarray(queries_prim, np.uint32)
    qb = np.ascontiguousarray(queries_b, np.float32).reshape(-1, 2)
    g = np.ascontiguousarray(np.broadcast_to(np.asarray(ng, np.float32), (len(qp), 3)))
    m = len(qp)
    outs = [np.empty((m, 3), np.float32) for _ in range(3)] + [np.empty((m, 2), np.float32),
                                                              np.empty((m, 3), np.float32)]
    rc = lib().apg_debug_shading(_ptr(tri), len(tri), nv, C.byref(sh), nt_,
                                 None if dn is None else _ptr(dn), None if dt is None else _ptr(dt),
                                 None if w is None else _ptr(w), C.byref(mat), _ptr(texels), _ptr(sizes),
                                 len(ts), _ptr(qp), _ptr(qb), _ptr(g), m, *[_ptr(o) for o in outs])
    if rc != APG_OK:
        raise ApgError(rc, "apg_debug_shading failed")
    return outs


def debug_triangles(v9, org, dir, tmax=None, watertight=False):
    """Host build of the kernels' triangle tests, brute force over all triangles (test hook):
    (prim, t, b0, b1) of the closest hit per ray."""
    v = np.ascontiguousarray(v9, np.float32).reshape(-1, 9)
    o = np.ascontiguousarray(org, np.float32).reshape(-1, 3)
    d = np.ascontiguousarray(dir, np.float32).reshape(-1, 3)
    n = len(o)
    tm = None if tmax is None else np.ascontiguousarray(np.broadcast_to(np.asarray(tmax, np.float32), (n,)))
    prim = np.empty(n, np.uint32)
    t, b0, b1 = (np.empty(n, np.float32) for _ in range(3))
    rc = lib().apg_debug_triangles(_ptr(v), len(v), _ptr(o), _ptr(d), None if tm is None else _ptr(tm), n,
                                   1 if watertight else 0, _ptr(prim), _ptr(t), _ptr(b0), _ptr(b1))
    if rc != APG_OK:
        raise ApgError(rc, "apg_debug_triangles failed")
    return prim, t, b0, b1


def debug_environment(rgb, dirs):
    e = np.ascontiguousarray(rgb, np.float32)
    d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    out = np.empty_like(d)
    rc = lib().apg_debug_environment(_ptr(e), e.shape[1], e.shape[0], _ptr(d), len(d), _ptr(out))
    if rc != APG_OK:
        raise ApgError(rc, "apg_debug_environment failed")
    return out
