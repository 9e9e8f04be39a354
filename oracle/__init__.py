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
