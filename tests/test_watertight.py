"""The watertight triangle test (Woop, Benthin and Wald, JCGT 2013), host build of the
kernels' code (traverse_core.h) through apg_debug_triangles.

The reference has no watertight test (its Moller-Trumbore, bvh_common.cpp:66-103, is
the default mode and is pinned elsewhere), so this mode is pinned by what the
mathematics fixes:
- a closed mesh seen from inside: every ray hits (rays aimed at shared vertices and
  edges are the cases a non-watertight test drops; Moller-Trumbore's misses are counted
  to show the cases are hard ones);
- float64 ray/plane solves: t and the barycentrics of every hit;
- the barycentric convention (b0 weights v1, b1 weights v2) and the closest hit;
- generic rays: the same triangle as Moller-Trumbore, t within float rounding;
- edges inclusive, both faces hit, t < tmax exclusive, t >= 0.
"""
import numpy as np
import pytest

import alpine_amd as A
from alpine_amd import scenes

MISS = 0xFFFFFFFF
T_REL_TOL = 2e-6     # float32 test against float64 solve, generic rays
B_ABS_TOL = 4e-6


def unit_sphere(levels=3):
    v, f = scenes.icosahedron()
    for _ in range(levels):
        v, f = scenes.subdivide(v, f)
    v = (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)
    return v, f.astype(np.uint32)


def inside_rays_at_vertices(v, n, seed):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-0.5, 0.5, (n, 3)).astype(np.float32)
    d = (v[rng.integers(0, len(v), n)] - o).astype(np.float32)
    return o, d


def inside_rays_at_edges(v, f, n, seed):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-0.5, 0.5, (n, 3)).astype(np.float32)
    e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    ei = e[rng.integers(0, len(e), n)]
    s = rng.uniform(0, 1, (n, 1)).astype(np.float32)
    tgt = (v[ei[:, 0]] + s * (v[ei[:, 1]] - v[ei[:, 0]])).astype(np.float32)
    return o, (tgt - o).astype(np.float32)


def solve64(v9, prim, o, d):
    """(t, b0, b1) of the ray / triangle-plane intersection in float64:
    o + t d = v0 + b0 (v1 - v0) + b1 (v2 - v0)."""
    V = v9[prim].reshape(-1, 3, 3).astype(np.float64)
    D = d.astype(np.float64)
    M = np.stack([-D, V[:, 1] - V[:, 0], V[:, 2] - V[:, 0]], axis=2)
    x = np.linalg.solve(M, (o.astype(np.float64) - V[:, 0])[..., None])[..., 0]
    return x[:, 0], x[:, 1], x[:, 2]


@pytest.mark.parametrize("aim", ["vertices", "edges"])
def test_closed_mesh_has_no_leaks(aim):
    v, f = unit_sphere(3)
    v9 = v[f].reshape(-1, 9)
    if aim == "vertices":
        o, d = inside_rays_at_vertices(v, 20_000, seed=3)
    else:
        o, d = inside_rays_at_edges(v, f, 20_000, seed=4)
    pw, tw, b0, b1 = A.debug_triangles(v9, o, d, watertight=True)
    pm = A.debug_triangles(v9, o, d, watertight=False)[0]
    mt_miss = int((pm == MISS).sum())
    print(f"rays aimed at {aim}: Moller-Trumbore misses {mt_miss} of {len(o)}, watertight 0 expected")
    assert (pw != MISS).all()
    assert mt_miss > 0   # the cases are the hard ones
    # every hit lies on the unit sphere's mesh, in front of the origin
    p = o.astype(np.float64) + tw[:, None].astype(np.float64) * d.astype(np.float64)
    r = np.linalg.norm(p, axis=1)
    assert (tw > 0).all() and r.max() <= 1.0 + 1e-6 and r.min() > 0.98
    assert (b0 >= 0).all() and (b1 >= 0).all() and (b0 + b1 <= 1 + 1e-6).all()


def test_hits_match_float64_solve_and_convention():
    v, f = unit_sphere(3)
    v9 = v[f].reshape(-1, 9)
    rng = np.random.default_rng(5)
    n = 20_000
    o = rng.uniform(-0.5, 0.5, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    p, t, b0, b1 = A.debug_triangles(v9, o, d, watertight=True)
    assert (p != MISS).all()
    t64, u64, w64 = solve64(v9, p, o, d)
    assert np.max(np.abs(t - t64) / t64) <= T_REL_TOL
    assert np.max(np.abs(b0 - u64)) <= B_ABS_TOL
    assert np.max(np.abs(b1 - w64)) <= B_ABS_TOL
    # hit point from the barycentrics (b0 weights v1, b1 weights v2) equals o + t d
    V = v9[p].reshape(-1, 3, 3).astype(np.float64)
    ph = (1 - b0 - b1)[:, None] * V[:, 0] + b0[:, None] * V[:, 1] + b1[:, None] * V[:, 2]
    pr = o + t[:, None].astype(np.float64) * d
    assert np.max(np.linalg.norm(ph - pr, axis=1)) <= 1e-5
    # closest: no other triangle's float64 plane hit inside it lies nearer by more than rounding
    pm, tm, _, _ = A.debug_triangles(v9, o, d, watertight=False)
    same = pm == p
    assert same.mean() > 0.999
    assert np.max(np.abs(tm[same] - t[same]) / t[same]) <= T_REL_TOL


def test_brute_force_closest_on_random_soup():
    """Overlapping random triangles: the closest watertight hit is the nearest float64
    plane hit among the triangles that contain the float64 hit point well inside."""
    v9 = scenes.random_triangles(300, seed=11, extent=1.0, lmin=0.2, lmax=0.6)
    v9 = v9[0][v9[1]].reshape(-1, 9).astype(np.float32)
    rng = np.random.default_rng(12)
    n = 3000
    o = rng.uniform(-1.5, 1.5, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    p, t, b0, b1 = A.debug_triangles(v9, o, d, watertight=True)
    m = len(v9)
    best = np.full(n, np.inf)
    for k in range(m):
        tk, uk, wk = solve64(v9, np.full(n, k), o, d)
        inside = (uk > 1e-4) & (wk > 1e-4) & (uk + wk < 1 - 1e-4) & (tk > 1e-4)
        best = np.where(inside & (tk < best), tk, best)
    found = np.isfinite(best)
    assert (p[found] != MISS).all()
    assert np.max(np.abs(t[found] - best[found]) / best[found]) <= 1e-5
    assert ((p != MISS) >= found).all()


def test_edges_inclusive_faces_and_tmax():
    # two triangles sharing the edge y = 0 in the plane z = 0
    v9 = np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0],
                   [0, 0, 0, 0, -1, 0, 1, 0, 0]], np.float32)
    o = np.array([[0.5, 0.0, 1.0],     # on the shared edge, exact zero edge function
                  [0.0, 0.0, 1.0],     # the shared vertex
                  [0.25, 0.25, -1.0],  # from below: the back face is hit too
                  [0.25, 0.25, 1.0],   # tmax == t: not a hit (t < tmax)
                  [0.25, 0.25, -1.0],  # pointing away: t < 0, no hit
                  [2.0, 2.0, 1.0]], np.float32)
    d = np.array([[0, 0, -1], [0, 0, -1], [0, 0, 1], [0, 0, -1], [0, 0, -1], [0, 0, -1]],
                 np.float32)
    tmax = np.array([np.inf, np.inf, np.inf, 1.0, np.inf, np.inf], np.float32)
    p, t, b0, b1 = A.debug_triangles(v9, o, d, tmax=tmax, watertight=True)
    assert p[0] in (0, 1) and t[0] == 1.0
    assert p[1] in (0, 1) and t[1] == 1.0
    assert p[2] == 0 and t[2] == 1.0 and b0[2] == 0.25 and b1[2] == 0.25
    assert p[3] == MISS and p[4] == MISS and p[5] == MISS
    # tmax just above t accepts
    tmax[3] = np.nextafter(np.float32(1.0), np.float32(2.0))
    assert A.debug_triangles(v9, o[3:4], d[3:4], tmax=tmax[3:4], watertight=True)[0][0] == 0


def test_degenerate_triangles_never_hit():
    v9 = np.array([[0, 0, 0, 1, 0, 0, 2, 0, 0],      # collinear
                   [0, 0, 0, 0, 0, 0, 0, 0, 0]], np.float32)
    rng = np.random.default_rng(2)
    o = rng.uniform(-1, 1, (500, 3)).astype(np.float32)
    d = -o + rng.normal(scale=1e-3, size=o.shape).astype(np.float32)
    assert (A.debug_triangles(v9, o, d, watertight=True)[0] == MISS).all()
