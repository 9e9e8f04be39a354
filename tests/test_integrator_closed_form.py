"""The integrator (alpine.cpp:290-432) pinned by what the mathematics fixes, on the oracle
(CPU) and on the GPU path tracer.

Each scene in tests/analytic_scenes.py has a closed-form pixel value: an open plane under a
uniform background (with and without tilted shading normals), a disk light on the camera's
axis (direct light by MIS, by light sampling alone and by BSDF sampling alone), a point light
(inverse square), quad and triangle lights (Lambert's polygon formula), a closed furnace
sphere (a geometric series in the albedo, bounded per sample), and Russian roulette on the
furnace (same expectation).  Planted mistakes in the oracle (a wrong MIS weight, no kill of
continuations below the geometric surface, shadow rays offset along ng, no throughput
update) must each fail one of the closed forms.

Tolerances: zero-variance cases 1e-4 relative (float rounding plus the narrow cameras'
1e-4 off-axis change); Monte Carlo cases 4 standard errors of the image mean (estimated from
the per-pixel means) plus the same 1e-4.
"""
import numpy as np
import pytest

import oracle
from tests import analytic_scenes as A


def _mean_err(img):
    """Image mean per channel and its standard error from the per-pixel means."""
    px = img.reshape(-1, 3)
    return px.mean(axis=0), px.std(axis=0, ddof=1) / np.sqrt(len(px))


def _close_mc(img, expect, rel=1e-4, k=4.0):
    m, se = _mean_err(img)
    tol = k * se + rel * np.abs(expect) + 1e-12
    return np.all(np.abs(m - expect) <= tol), m, tol


def _close_exact(img, expect, rel=1e-4):
    return np.allclose(img, np.broadcast_to(expect, img.shape), rtol=rel, atol=1e-7)


# ---- CPU: the oracle ---------------------------------------------------------------------
@pytest.mark.parametrize("depth", [1, 2, 5])
def test_oracle_open_plane_is_rho_times_background(depth):
    c = A.open_plane(depth)
    img = oracle_img = A.oracle_image(c)
    assert _close_exact(oracle_img, c.expect, 1e-5), (img.reshape(-1, 3)[:3], c.expect)


def test_oracle_open_plane_tilted_normals_kills_below_ng():
    c = A.open_plane(3, theta=np.radians(60.0), spp=32)
    ok, m, tol = _close_mc(A.oracle_image(c), c.expect)
    assert ok, (m, c.expect, tol)
    # without the dot(wi, ng) <= 0 kill the killed quarter keeps bouncing off the plane
    ok, m, _ = _close_mc(A.oracle_image(c, plant=oracle.PLANT_NO_NG_KILL), c.expect)
    assert not ok, ("planted missing kill not caught", m, c.expect)


@pytest.mark.parametrize("mode", [oracle.MIS_BOTH, oracle.MIS_LIGHT_ONLY, oracle.MIS_BSDF_ONLY])
@pytest.mark.parametrize("theta", [0.0, np.radians(30.0)])
def test_oracle_disk_irradiance_three_strategies(mode, theta):
    c = A.disk_on_axis(theta)
    ok, m, tol = _close_mc(A.oracle_image(c, mis_mode=mode), c.expect)
    assert ok, (mode, m, c.expect, tol)


def test_oracle_wrong_mis_weight_is_caught():
    c = A.disk_on_axis()
    ok, m, _ = _close_mc(A.oracle_image(c, plant=oracle.PLANT_MIS_WEIGHT), c.expect)
    assert not ok and np.all(m > c.expect), ("planted MIS weight not caught", m, c.expect)


def test_oracle_point_light_inverse_square():
    c = A.point_above()
    assert _close_exact(A.oracle_image(c), c.expect)


@pytest.mark.parametrize("kind", [A.QUAD, A.TRIANGLE])
def test_oracle_polygon_light_lambert_formula(kind):
    c = A.polygon_above(kind)
    ok, m, tol = _close_mc(A.oracle_image(c), c.expect)
    assert ok, (kind, m, c.expect, tol)


def test_oracle_shadow_rays_offset_along_shading_normal():
    c = A.grazing_point_with_ledge()
    assert np.all(A.oracle_image(c) == 0.0), "the ledge shadows every pixel"
    img = A.oracle_image(c, plant=oracle.PLANT_SHADOW_NG)
    assert _close_exact(img, c.unshadowed, 1e-3), (img.reshape(-1, 3)[:2], c.unshadowed)


def _furnace_ok(px, c, spp):
    """Every pixel inside the closed-form bounds, except that the reference's Moller-Trumbore
    test is not watertight: a ray through a shared edge can leave the sphere (and its sample
    then ends early), which a few pixels of a large image may show.  Such a pixel is still
    above (spp - 1) / spp of the lower bound, at most 1 in 1,000 pixels may have one, and none
    may exceed the upper bound."""
    hi_ok = np.all(px <= c.hi * (1 + 1e-5))
    low = np.any(px < c.lo * (1 - 1e-5), axis=1)
    leak_ok = np.all(px[low] >= c.lo * (spp - 1) / spp * (1 - 1e-5)) and low.mean() <= 1e-3
    return hi_ok and leak_ok, int(low.sum())


@pytest.mark.parametrize("depth,rho", [(1, (0.5, 0.5, 0.5)), (2, (0.8, 0.5, 0.3)), (5, (0.9, 0.6, 0.2))])
def test_oracle_furnace_geometric_series(depth, rho):
    c = A.furnace(depth, rho, subdiv=4, spp=4)
    img = A.oracle_image(c)
    px = img.reshape(-1, 3)
    ok, nlow = _furnace_ok(px, c, 4)
    assert ok, (nlow, px.min(0), px.max(0), c.lo, c.hi)
    if depth > 1:
        bad = A.oracle_image(c, plant=oracle.PLANT_NO_THROUGHPUT).reshape(-1, 3)
        assert np.all(bad > c.hi), "planted missing throughput not caught"


def test_oracle_russian_roulette_keeps_the_mean():
    c = A.furnace(8, (0.8, 0.6, 0.4), rr_depth=2, subdiv=4, spp=16)
    img = A.oracle_image(c)
    m, se = _mean_err(img)
    assert np.all(m >= c.lo - 4 * se) and np.all(m <= c.hi + 4 * se), (m, se, c.lo, c.hi)
    assert np.all(img.reshape(-1, 3).std(axis=0) > 0), "roulette adds variance"


# ---- GPU: the same scenes through the wavefront path tracer -------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("depth", [1, 2, 5])
def test_gpu_open_plane_is_rho_times_background(depth):
    print(f"[gpu] closed form: open plane depth {depth}", flush=True)
    c = A.open_plane(depth)
    assert _close_exact(A.gpu_image(c), c.expect, 1e-5)


@pytest.mark.gpu
def test_gpu_open_plane_tilted_normals():
    print("[gpu] closed form: tilted shading normals", flush=True)
    c = A.open_plane(3, theta=np.radians(60.0), spp=32)
    ok, m, tol = _close_mc(A.gpu_image(c), c.expect)
    assert ok, (m, c.expect, tol)


@pytest.mark.gpu
@pytest.mark.parametrize("theta", [0.0, np.radians(30.0)])
def test_gpu_disk_irradiance(theta):
    print(f"[gpu] closed form: disk light theta {theta:.3f}", flush=True)
    c = A.disk_on_axis(theta)
    ok, m, tol = _close_mc(A.gpu_image(c), c.expect)
    assert ok, (m, c.expect, tol)


@pytest.mark.gpu
def test_gpu_point_light_inverse_square():
    print("[gpu] closed form: point light", flush=True)
    c = A.point_above()
    assert _close_exact(A.gpu_image(c), c.expect)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", [A.QUAD, A.TRIANGLE])
def test_gpu_polygon_light_lambert_formula(kind):
    print(f"[gpu] closed form: polygon light kind {kind}", flush=True)
    c = A.polygon_above(kind)
    ok, m, tol = _close_mc(A.gpu_image(c), c.expect)
    assert ok, (kind, m, c.expect, tol)


@pytest.mark.gpu
def test_gpu_shadow_rays_offset_along_shading_normal():
    print("[gpu] closed form: shadow-ray offset", flush=True)
    c = A.grazing_point_with_ledge()
    assert np.all(A.gpu_image(c) == 0.0)


@pytest.mark.gpu
@pytest.mark.parametrize("depth,rho", [(1, (0.5, 0.5, 0.5)), (2, (0.8, 0.5, 0.3)), (5, (0.9, 0.6, 0.2))])
def test_gpu_furnace_geometric_series(depth, rho):
    print(f"[gpu] closed form: furnace depth {depth}", flush=True)
    c = A.furnace(depth, rho, spp=16)
    # the watertight test loses no ray: every pixel inside the bounds
    px = A.gpu_image(c, slots=4, watertight=True).reshape(-1, 3)
    assert np.all(px >= c.lo * (1 - 1e-5)) and np.all(px <= c.hi * (1 + 1e-5)), (px.min(0), px.max(0), c.lo, c.hi)
    # the reference's test: the same, up to the rare leaked sample
    px = A.gpu_image(c, slots=4).reshape(-1, 3)
    ok, nlow = _furnace_ok(px, c, 16)
    assert ok, (nlow, px.min(0), px.max(0), c.lo, c.hi)


@pytest.mark.gpu
def test_gpu_russian_roulette_keeps_the_mean():
    """Same furnace at depth 8 with and without roulette from depth 2 (the option configs 3
    and 4 run with): both means inside the closed-form bounds within 4 standard errors."""
    print("[gpu] closed form: Russian roulette", flush=True)
    c = A.furnace(8, (0.8, 0.6, 0.4), rr_depth=2, spp=64)
    for rr in (2, 0):
        img = A.gpu_image(c, slots=4, rr_depth=rr)
        m, se = _mean_err(img)
        assert np.all(m >= c.lo - 4 * se) and np.all(m <= c.hi + 4 * se), (rr, m, se, c.lo, c.hi)
