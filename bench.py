#!/usr/bin/env python3
"""Traversal microbench on one MI355X: Mrays/s for incoherent closest-hit rays.

Workload (one "step" = one closest-hit pass over all rays, inputs resident in HBM):
  displaced icosphere (subdivision 8, radius 1 + 0.1 * value noise, seed 7) plus
  200,000 random triangles -> 1,510,720 fp32 triangles in a quantised 8-wide BVH;
  4,194,304 incoherent rays (origins uniform in the scene bounds, directions
  uniform on the sphere), tmax = FLT_MAX.

Also reported (not the headline value): coherent closest-hit (2048 x 2048
pinhole), any-hit (tmax = FLT_MAX and tmax ~ U(0, |diag|)), build time, SAH
cost, the roofline of the traversal kernel (algorithmic bytes per ray from an
instrumented run of the same kernel: 28 B ray + 16 B hit + 64 B per node fetch +
36 B per triangle fetch) and the reference CPU path timed on a bounded ray
sample on this host's cores.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s incoherent closest-hit; time to equal relMSE, guided vs plain GPU PT"
HBM_PEAK_GBS = 8000.0
N_RAYS = 4_194_304
B_RAY_IN, B_HIT_OUT, B_OCC_OUT, B_NODE, B_TRI = 28, 16, 4, 64, 36
# binned-SAH leaf threshold of the microbench tree: one triangle per leaf where the SAH allows
# (a triangle test through the wave's queue costs more than the node visits it saves)
LEAF_THRESHOLD = 1
# the animated config rebuilds every frame: leaf threshold 2 builds its forest 7 ms faster for
# the same render time
BATCH_LEAF_THRESHOLD = 2
EAGER_FRAMES = 5  # config 4 frames run stage by stage for the rebuild / render split


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=262_144)
    ap.add_argument("--builder", choices=["gpu", "host"], default="gpu")
    ap.add_argument("--no-render", action="store_true", help="skip the path-tracing configs")
    ap.add_argument("--frames", type=int, default=60,
                    help="frames of the batched animated config (the workload states 60)")
    ap.add_argument("--autotune", action="store_true",
                    help="sweep the launch shapes first (the default; kept for old command lines)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the library's default launch shapes (no sweeps)")
    ap.add_argument("--trace-launch", metavar="B,P,F,S,T",
                    help="use this traversal launch shape (block size, blocks per CU, fetch "
                         "threshold, LDS stack entries, triangle batch) instead of the sweep, "
                         "e.g. to profile exactly the shape a bench run picked")
    ap.add_argument("--render-config", choices=sorted(RENDER_CONFIGS),
                    help="run only this path-tracing config (or the MLP kernels) and print its "
                         "figures as one JSON line; for profiling, not the headline measurement")
    return ap.parse_args()


def timed_passes(fn, steps, warmup):
    """Run fn() warmup + steps times; return per-step ms (HIP events on torch's stream)."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    stop = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    t0 = time.perf_counter()
    for i in range(steps):
        start[i].record()
        fn()
        stop[i].record()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    ev = [s.elapsed_time(e) for s, e in zip(start, stop)]
    return float(np.mean(ev)), wall


PROFILE_JSON = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                            "trace_kernel_incoherent_closest_1p5Mtri_4Mrays.json")


def _profiled_traffic():
    """HBM-side bytes per launch of this kernel from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE runs (each scaled by the counter's reading for a 4 GiB copy), or None."""
    try:
        with open(PROFILE_JSON) as f:
            return int(json.load(f)["traffic_bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


def per_ray_host_rate(acc, org, dir, n=1 << 18, threads=256):
    """The drop-in adapter's per-ray path: `threads` host threads (alpine's 256 tile tasks,
    alpine.cpp:276-285) each calling apg_intersect_host / apg_occluded_host with one ray per
    call over the first n incoherent rays (host memory in and out, coalesced into shared
    launches inside the library)."""
    rays6 = np.ascontiguousarray(np.concatenate([org[:n], dir[:n]], 1), np.float32)
    acc.host_rate(rays6[:4096], threads)  # warm-up
    _, s_c = acc.host_rate(rays6, threads)
    _, s_a = acc.host_rate(rays6, threads, any_hit=True)
    # the same threads handing over 16 rays per call (a tile row per call)
    _, s_16 = acc.host_rate(rays6, threads, rays_per_call=16)
    return {"threads": threads, "rays": n, "rays_per_call": 1,
            "closest_mrays": round(n / s_c / 1e6, 3), "any_mrays": round(n / s_a / 1e6, 3),
            "closest_mrays_16_per_call": round(n / s_16 / 1e6, 3)}


def hbm_read_gbs(gib=4, reps=5):
    """Achievable HBM read bandwidth: the library's streaming-read kernel (16-byte
    non-temporal loads, 8 workgroups of 256 per CU) over `gib` GiB, HIP events on its stream."""
    import ctypes as C
    import torch
    import alpine_amd as aa
    n = gib << 30
    x = torch.ones(n // 4, dtype=torch.float32, device="cuda")
    sink = torch.zeros(1, dtype=torch.int32, device="cuda")
    blocks = torch.cuda.get_device_properties(0).multi_processor_count * 8
    s = torch.cuda.current_stream()
    call = lambda: aa.lib().apg_debug_stream_read(C.c_void_p(x.data_ptr()), n, C.c_void_p(sink.data_ptr()),
                                                  blocks, C.c_void_p(s.cuda_stream))
    for _ in range(2):
        assert call() == 0
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        call()
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / reps
    del x
    return n / (ms * 1e-3) / 1e9


def hbm_copy_gbs():
    """Device-to-device copy of 4 GiB (read + write bytes), for comparison."""
    import torch
    n = 1 << 30
    x = torch.empty(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    x.fill_(1.0)
    for _ in range(2):
        y.copy_(x)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(5):
        y.copy_(x)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / 5
    del x, y
    return 2 * n * 4 / (ms * 1e-3) / 1e9


def _render_setup(sc, builder):
    import alpine_amd as aa
    acc = aa.GpuAccelerator(0)
    acc.append_mesh(sc.verts, sc.tris)
    acc.update_scene(builder=builder)
    acc.set_materials([aa.material(*m) for m in sc.materials], sc.tri_mat)
    acc.set_lights([aa.light(**l) for l in sc.lights], sc.light_sampler)
    acc.set_camera(aa.camera(**sc.camera))
    return acc


def _timed_render(acc, sc, spp, seed, slots=1, **kw):
    import torch
    acc.render_reset(sc.width, sc.height, seed, slots=slots)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    acc.render(spp, sc.max_depth, sc.rr_depth, sc.background, **kw)
    img, img64, total = acc.render_read(fp64=kw.get("fp64", False))
    return time.perf_counter() - t0, img, img64


def render_benchmarks(builder, frames=60, tune=True):
    """Path-tracing configs: Cornell room (16 spp, 256^2), many-light GGX (1920x1080, 64 spp,
    light BVH, RR from depth 3), 32 animated scenes in one tiled frame, guided vs plain
    time to equal relMSE (box grid behind a doorway) and the guide network's kernels."""
    out = {}
    out.update(cornell_benchmark(builder))
    log("[bench] many-light scene")
    out.update(many_light_benchmark(builder, tune=tune))
    log("[bench] batched animated scenes")
    out["batched_32_scenes_512_32spp"] = batched_benchmark(frames=frames, tune=tune)
    log("[bench] guided vs plain")
    out["guided_vs_plain"] = guided_benchmark(builder)
    log("[bench] guide network kernels")
    out["guide_mlp"] = mlp_benchmark()
    return out


def cornell_benchmark(builder):
    """Config 1: the Cornell room, 256 x 256, 16 spp, against a 1024-spp FP64 reference."""
    from alpine_amd import scenes
    out = {}
    sc = scenes.cornell_scene()
    acc = _render_setup(sc, builder)
    _timed_render(acc, sc, 1, 0)
    dt, img, _ = _timed_render(acc, sc, sc.spp, sc.seed)
    c = acc.render_counters()
    ref_spp = 1024
    _, _, ref64 = _timed_render(acc, sc, ref_spp, 1000, fp64=True)
    ref = ref64 / ref_spp
    out["cornell_256_16spp"] = {
        "seconds": round(dt, 4), "samples_per_s": round(sc.width * sc.height * sc.spp / dt, 1),
        "rel_mse_vs_1024spp_fp64": round(scenes.rel_mse(img / sc.spp, ref), 6),
        "rays_per_s": round((c["closest_rays"] + c["shadow_rays"]) / dt, 1) if dt > 0 else None}
    # the same 16 spp traced side by side (16 slots per pixel, one pass)
    import torch
    acc.render_reset(sc.width, sc.height, sc.seed, slots=16)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    acc.render(sc.spp, sc.max_depth, sc.rr_depth, sc.background)
    img, _, _ = acc.render_read()
    dt16 = time.perf_counter() - t0
    out["cornell_256_16spp"]["slots16"] = {
        "seconds": round(dt16, 4), "samples_per_s": round(sc.width * sc.height * sc.spp / dt16, 1),
        "rel_mse_vs_1024spp_fp64": round(scenes.rel_mse(img / sc.spp, ref), 6)}
    return out


def many_light_benchmark(builder, tune=True):
    """Config 3: 1920 x 1080, 64 spp, GGX materials, 10,000 emissive triangles behind the
    light BVH, RR from depth 3; relMSE against a 1024-spp FP64 reference.  With tune, the
    render stages' grids are swept on 4-spp passes of the same frame first."""
    from alpine_amd import scenes
    out = {}
    sc = scenes.many_light_scene()
    acc = _render_setup(sc, builder)
    launch = None
    sl = 1
    if tune:
        from alpine_amd import autotune
        # samples in flight per pixel too: the late bounces of a pass hold few paths, so more
        # samples per pass keep those launches fuller
        launch, _ = autotune.autotune_render(acc, sc.width, sc.height, max(ML_SLOTS), sc.max_depth,
                                             sc.rr_depth, sc.background, slots=ML_SLOTS, log=log)
        sl = launch["slots"]
    _timed_render(acc, sc, sl, 0, slots=sl)
    dt, img, _ = _timed_render(acc, sc, sc.spp, sc.seed, slots=sl)
    c = acc.render_counters()
    # convergence: the same estimator at 16 spp and a 1024-spp FP64 reference (other seeds)
    _, img16, _ = _timed_render(acc, sc, 16, sc.seed + 7, slots=sl)
    _, _, ref64 = _timed_render(acc, sc, ML_REF_SPP, 1000, slots=sl, fp64=True, graph=True)
    ref = ref64 / ML_REF_SPP
    bps = render_bytes_per_sample(acc, sc)
    sps = sc.width * sc.height * sc.spp / dt
    out["many_light_1080p_64spp"] = {
        "triangles": int(len(sc.tris)), "lights": len(sc.lights), "seconds": round(dt, 3),
        "samples_per_s": round(sc.width * sc.height * sc.spp / dt, 1),
        "rays_per_s": round((c["closest_rays"] + c["shadow_rays"]) / dt, 1),
        "finite": bool(np.isfinite(img).all()),
        "rel_mse_vs_1024spp_fp64": round(scenes.rel_mse(img / sc.spp, ref), 6),
        "rel_mse_16spp_vs_1024spp_fp64": round(scenes.rel_mse(img16 / 16, ref), 6),
        "bytes_per_sample": bps, "samples_frac_hbm": round(bps["total"] * sps / 1e9 / HBM_PEAK_GBS, 4),
        "launch": launch or acc.render_launch()}
    return out


# Algorithmic queue and path-state bytes per shaded path vertex: the shade stage reads the ray
# (origin, direction, path index: 28 B), its hit (16 B) and the path state (sampler 16 B,
# throughput 12 B, radiance 12 B) and writes the continuation ray (28 B), the path state (40 B)
# and two shadow rays (origin, direction, tmax, path index: 2 x 32 B); the two NEE terms
# (2 x 12 B) are written by shade and read by resolve, with the two occlusion bytes.
B_PATH_VERTEX = (28 + 16 + 40) + (28 + 40 + 2 * 32) + 2 * (2 * 12) + 2


def render_bytes_per_sample(acc, sc, spp=4, seed=77):
    """Algorithmic HBM bytes per sample of a render config, from one pass of the instrumented
    traversal: closest rays 28 B in + 16 B hit, shadow rays 28 B + 4 B result, 64 B per node and
    36 B per triangle fetched, plus B_PATH_VERTEX per shaded vertex."""
    acc.render_reset(sc.width, sc.height, seed)
    acc.render(spp, sc.max_depth, sc.rr_depth, sc.background, count_fetches=True)
    c, f = acc.render_counters(), acc.render_fetch_counts()
    n = sc.width * sc.height * spp
    trav = (c["closest_rays"] * (B_RAY_IN + B_HIT_OUT) + c["shadow_rays"] * (B_RAY_IN + B_OCC_OUT)
            + B_NODE * (f["closest_nodes"] + f["shadow_nodes"]) + B_TRI * (f["closest_tris"] + f["shadow_tris"]))
    queue = c["vertices"] * B_PATH_VERTEX
    return {"total": round((trav + queue) / n, 1), "traversal": round(trav / n, 1), "queues": round(queue / n, 1),
            "closest_rays": round(c["closest_rays"] / n, 3), "shadow_rays": round(c["shadow_rays"] / n, 3),
            "nodes_per_closest_ray": round(f["closest_nodes"] / max(c["closest_rays"], 1), 2),
            "nodes_per_shadow_ray": round(f["shadow_nodes"] / max(c["shadow_rays"], 1), 2)}


# The guide network: 96 -> 64 -> 64 -> 64 -> 64 -> 48 (bias adds not counted).  Forward FLOPs
# per query are 2 * sum(fan_in * fan_out); a training sample adds the backward pass (input and
# weight gradients, twice the forward), 3x in all.
MLP_LAYERS = [(96, 64), (64, 64), (64, 64), (64, 64), (64, 48)]
MLP_FWD_FLOP = 2 * sum(a * b for a, b in MLP_LAYERS)
MLP_TRAIN_FLOP = 3 * MLP_FWD_FLOP
BF16_MFMA_PEAK_TFLOPS = 2500.0


def mlp_benchmark(n_infer=1 << 20, n_train=1 << 19, reps=10):
    """The guide network's fused kernels on their own: inference (encoding fused in) on
    n_infer queries and training steps (encoding, forward, loss head, backward, Adam) on
    n_train samples, as in a training pass; TFLOP/s against the dense BF16 MFMA peak."""
    import torch
    import alpine_amd as aa
    acc = aa.GpuAccelerator(0)
    acc.guide_init(seed=1, lr=1e-3)
    g = torch.Generator(device="cpu").manual_seed(3)

    def queries(n):
        p = torch.rand(n, 3, generator=g)
        nn = torch.nn.functional.normalize(torch.randn(n, 3, generator=g), dim=1)
        wo = torch.nn.functional.normalize(torch.randn(n, 3, generator=g), dim=1)
        return torch.cat([p, nn, wo], 1).contiguous().cuda()

    q = queries(n_infer)
    ms_inf, _ = timed_passes(lambda: acc.guide_infer(q), reps, 2)
    qt = queries(n_train)
    dirs = torch.nn.functional.normalize(torch.randn(n_train, 3, generator=g), dim=1).cuda()
    w = torch.rand(n_train, generator=g).cuda()
    loss0, _ = acc.guide_train_queries(qt, dirs, w)
    # asynchronous steps (gradient kernel, reduction + Adam, weight images), HIP events
    ms_tr, wall_tr = timed_passes(lambda: acc.guide_train_queries(qt, dirs, w, want_loss=False), reps, 2)
    loss1, _ = acc.guide_train_queries(qt, dirs, w)
    tf_inf = n_infer * MLP_FWD_FLOP / (ms_inf * 1e-3) / 1e12
    tf_tr = n_train * MLP_TRAIN_FLOP / (ms_tr * 1e-3) / 1e12
    return {
        "layers": MLP_LAYERS, "forward_flop_per_query": MLP_FWD_FLOP,
        "train_flop_per_sample": MLP_TRAIN_FLOP,
        "infer_queries": n_infer, "infer_ms": round(ms_inf, 4), "infer_tflops": round(tf_inf, 2),
        "infer_frac_bf16_peak": round(tf_inf / BF16_MFMA_PEAK_TFLOPS, 4),
        "train_samples": n_train, "train_step_ms": round(ms_tr, 4), "train_tflops": round(tf_tr, 2),
        "train_frac_bf16_peak": round(tf_tr / BF16_MFMA_PEAK_TFLOPS, 4),
        "train_step_wall_ms": round(wall_tr, 4),
        "train_loss_first_last": [round(loss0, 4), round(loss1, 4)],
        "note": "train step = gradient kernel + reduction/Adam + weight images, enqueued "
                "without a host round trip and timed with HIP events",
    }


# Guided-rendering protocol: plain PT renders PLAIN_SPP samples per pixel progressively; its
# final relMSE against a REF_SPP FP64 plain reference is the target.  The guided run trains
# online for TRAIN_SPP samples per pixel (these samples stay in the image) and then renders with
# the frozen proposal until it reaches the target.  Both run SLOTS samples of each pixel side by
# side and are timed around the GPU work only (the relMSE checks read the image between chunks).
GUIDE_PLAIN_SPP, GUIDE_REF_SPP, GUIDE_SLOTS, GUIDE_CHUNK = 1024, 16384, 4, 32
# FP64 references of the many-light frame and of the batched frame
ML_REF_SPP, BATCH_REF_SPP = 1024, 256
ML_SLOTS = (2, 4, 8, 16)  # samples in flight per pixel the many-light config's autotuning tries
BATCH_SLOTS = (2, 4)  # and the batched config's (4096 x 2048 paths per slot)
GUIDE_TRAIN_SPP, GUIDE_STEPS, GUIDE_FRAC, GUIDE_SAMPLES = 64, 64, 0.5, 1 << 19
# Adam learning rates: the workload's stated 1e-3 (the headline figure), and 5e-3, the best
# of the settings swept on this scene, reported beside it
GUIDE_LR, GUIDE_LR_TUNED = 1e-3, 5e-3
GUIDE_DEPTH = 0   # path vertices below this depth sample the proposal (0 = every vertex)
GUIDE_STEP_SPP = 8  # the guided run's relMSE is checked every 8 spp (two passes of 4 slots)


def guided_benchmark(builder):
    import torch
    from alpine_amd import scenes
    sc = scenes.guiding_scene()
    acc = _render_setup(sc, builder)

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    acc.render_reset(sc.width, sc.height, 0, slots=GUIDE_SLOTS)
    acc.render(GUIDE_SLOTS, sc.max_depth)
    acc.render_reset(sc.width, sc.height, 4000, slots=GUIDE_SLOTS)
    log(f"[bench] guided: {GUIDE_REF_SPP}-spp reference")
    t_ref = timed(lambda: acc.render(GUIDE_REF_SPP, sc.max_depth, fp64=True))
    _, r64, _ = acc.render_read(fp64=True)
    ref = r64 / GUIDE_REF_SPP
    # plain, progressive
    acc.render_reset(sc.width, sc.height, 11, slots=GUIDE_SLOTS)
    t_plain, spp = 0.0, 0
    curve_plain = []
    while spp < GUIDE_PLAIN_SPP:
        t_plain += timed(lambda: acc.render(GUIDE_CHUNK, sc.max_depth, graph=True))
        spp += GUIDE_CHUNK
        img, _, _ = acc.render_read()
        curve_plain.append((spp, t_plain, scenes.rel_mse(img / spp, ref)))
    target = curve_plain[-1][2]

    def guided(lr, seed):
        # online training, then the frozen proposal until the target is reached
        acc.guide_init(seed=1, lr=lr)
        acc.render_reset(sc.width, sc.height, seed, slots=GUIDE_SLOTS)
        t_train = timed(lambda: acc.render(GUIDE_TRAIN_SPP, sc.max_depth, guide=1, guide_warmup=1,
                                           train_steps=GUIDE_STEPS, guide_frac=GUIDE_FRAC,
                                           train_samples=GUIDE_SAMPLES))
        img, _, tot = acc.render_read()
        t_g = t_train
        curve_g = [(tot, t_g, scenes.rel_mse(img / tot, ref))]
        reach = t_g if curve_g[0][2] <= target else None
        while reach is None and tot < 4 * GUIDE_PLAIN_SPP:
            t_g += timed(lambda: acc.render(GUIDE_STEP_SPP, sc.max_depth, guide=2, guide_frac=GUIDE_FRAC,
                                            guide_depth=GUIDE_DEPTH, graph=True))
            img, _, tot = acc.render_read()
            e = scenes.rel_mse(img / tot, ref)
            curve_g.append((tot, t_g, e))
            if e <= target:
                reach = t_g
        # equal-time comparison: plain's relMSE at the guided run's wall time
        e_plain_eq = next((e for s_, t_, e in curve_plain if t_ >= t_g), curve_plain[-1][2])
        return {
            "lr": lr, "train_spp": GUIDE_TRAIN_SPP, "train_seconds": round(t_train, 4),
            "spp_at_target": int(tot), "seconds_to_target": None if reach is None else round(reach, 4),
            "rel_mse": round(curve_g[-1][2], 6),
            "time_ratio_guided_over_plain": None if reach is None else round(reach / t_plain, 4),
            "plain_rel_mse_at_guided_time": round(e_plain_eq, 6),
        }

    log(f"[bench] guided: plain reached relMSE {target:.5f} in {t_plain:.3f} s")
    stated = guided(GUIDE_LR, 21)
    log(f"[bench] guided: lr {GUIDE_LR}: {stated['time_ratio_guided_over_plain']} of plain's time")
    tuned = guided(GUIDE_LR_TUNED, 31)
    log(f"[bench] guided: lr {GUIDE_LR_TUNED}: {tuned['time_ratio_guided_over_plain']} of plain's time")
    return {
        "scene": "box grid (config-3 geometry) behind a 4 x 3 doorway, one 1 x 1 quad light",
        "triangles": int(len(sc.tris)), "frame": [sc.width, sc.height], "slots": GUIDE_SLOTS,
        "reference": f"{GUIDE_REF_SPP} spp plain, FP64 accumulation ({t_ref:.1f} s)",
        "plain_spp": GUIDE_PLAIN_SPP, "plain_seconds": round(t_plain, 4), "plain_rel_mse": round(target, 6),
        "time_ratio_guided_over_plain": stated["time_ratio_guided_over_plain"],
        "guided_lr_stated": stated, "guided_lr_tuned": tuned,
        "train": {"minibatches_per_pass": GUIDE_STEPS, "proposal_fraction": GUIDE_FRAC,
                  "samples_per_pass": GUIDE_SAMPLES},
    }


def batched_benchmark(frames=60, spp=32, max_depth=5, rr_depth=3, seed=1, tune=True):
    """Many independent scenes, animated, one tiled frame: 32 scenes of 10 K - 2 M triangles
    (log-uniform, about 12 M in all) in a per-shape forest, 512 x 512 tiles, 32 spp per frame.
    Each frame turns the last 25 % of every scene's triangles about y on the GPU, rebuilds all
    32 trees on the GPU in one forest build, and renders every tile in the same launches.  A few
    frames run stage by stage (the rebuild / render split); the stated `frames` frames run as
    one HIP graph each (apg_frame_render: transforms, rebuild, reset and render replayed)."""
    import torch
    import alpine_amd as aa
    from alpine_amd import scenes
    t0 = time.perf_counter()
    B = scenes.BatchedScenes()
    t_gen = time.perf_counter() - t0
    acc = aa.GpuAccelerator(0)
    for v, f in B.meshes:
        acc.append_mesh(v, f)
    t0 = time.perf_counter()
    acc.update_scene(leaf_threshold=BATCH_LEAF_THRESHOLD, per_shape=True, stats=False)
    torch.cuda.synchronize()
    t_first = time.perf_counter() - t0
    n_tris = B.triangles
    acc.set_materials([aa.material(*m) for m in B.materials], np.zeros(n_tris, np.uint32))
    acc.set_lights([aa.light(**l) for l in B.lights], 0)
    acc.set_tiles([aa.camera(**c) for c in B.cams], B.tile, B.tile, roots=np.arange(len(B.cams)))
    W, H = B.tile * B.tiles_x, B.tile * B.tiles_y
    launch = None
    sl = 1
    if tune:
        from alpine_amd import autotune
        launch, _ = autotune.autotune_render(acc, W, H, 4, max_depth, rr_depth, slots=BATCH_SLOTS, log=log,
                                             graph=True)
        sl = launch["slots"]
    rot = B.rotation(2.0 * np.pi / 120.0)
    xf = [(k, first, count, rot) for k, (first, count) in enumerate(B.anim_range)]
    # the stage split, from eager frames: transforms + rebuild, then reset + render
    builds, renders = [], []
    for f in range(EAGER_FRAMES):
        torch.cuda.synchronize()
        a = time.perf_counter()
        for k, first, count, m in xf:
            acc.transform_vertices(k, m, first=first, count=count)
        acc.update_scene(leaf_threshold=BATCH_LEAF_THRESHOLD, per_shape=True, stats=False)
        torch.cuda.synchronize()
        b = time.perf_counter()
        acc.render_reset(W, H, seed + f, slots=sl)
        acc.render(spp, max_depth, rr_depth, graph=True)
        torch.cuda.synchronize()
        c = time.perf_counter()
        builds.append(b - a)
        renders.append(c - b)
    cnt = acc.render_counters()
    # the stated frame loop: every frame (transforms, rebuild, reset, render) one HIP graph
    frame_s, short = [], 0
    for f in range(frames):
        torch.cuda.synchronize()
        a = time.perf_counter()
        acc.frame_render(xf, spp, max_depth, rr_depth, seed=seed + EAGER_FRAMES + f,
                         leaf_threshold=BATCH_LEAF_THRESHOLD, per_shape=True)
        short += acc.frame_status() != 0
        frame_s.append(time.perf_counter() - a)
    eager_frames, captures, replays = acc.frame_counts()
    img, _, total = acc.render_read()
    # convergence of the last frame: a BATCH_REF_SPP FP64 reference of the same geometry
    acc.render_reset(W, H, 5000, slots=sl)
    acc.render(BATCH_REF_SPP, max_depth, rr_depth, fp64=True, graph=True)
    _, ref64, _ = acc.render_read(fp64=True)
    rel = scenes.rel_mse(img / total, ref64 / BATCH_REF_SPP)
    # stage-by-stage frames again after the graph loop (clock drift over the run shows here)
    after = []
    for f in range(EAGER_FRAMES):
        torch.cuda.synchronize()
        a = time.perf_counter()
        for k, first, count, m in xf:
            acc.transform_vertices(k, m, first=first, count=count)
        acc.update_scene(leaf_threshold=BATCH_LEAF_THRESHOLD, per_shape=True, stats=False)
        acc.render_reset(W, H, seed + 1000 + f, slots=sl)
        acc.render(spp, max_depth, rr_depth, graph=True)
        torch.cuda.synchronize()
        after.append(time.perf_counter() - a)
    r = float(np.median(renders))
    fr = float(np.median(frame_s[1:]))  # the first graph frame is the eager one
    return {
        "scenes": len(B.meshes), "triangles": n_tris, "frame": [W, H], "spp": spp,
        "frames": frames, "max_depth": max_depth, "rr_depth": rr_depth,
        "generate_s": round(t_gen, 2), "first_build_upload_ms": round(t_first * 1e3, 1),
        "rebuild_ms_median": round(float(np.median(builds)) * 1e3, 1),
        "render_ms_median": round(r * 1e3, 1),
        "frame_ms_median_eager": round(float(np.median(np.array(builds) + np.array(renders))) * 1e3, 1),
        "frame_ms_median_eager_after": round(float(np.median(after)) * 1e3, 1),
        "frame_ms_median": round(fr * 1e3, 1),
        "frame_graph": {"frames": frames, "eager": eager_frames, "captures": captures, "replays": replays,
                        "short_rebuilds": int(short), "frame_ms_median": round(fr * 1e3, 1),
                        "frames_per_s": round(1.0 / fr, 2),
                        "samples_per_s": round(W * H * spp / fr, 1)},
        "samples_per_s": round(W * H * spp / r, 1),
        "rays_per_s": round((cnt["closest_rays"] + cnt["shadow_rays"]) / renders[-1], 1),
        "rebuild_mtris_per_s": round(n_tris / float(np.median(builds)) / 1e6, 2),
        "rebuild_frac_hbm": round(build_bytes(n_tris, acc.stats()["wide_nodes"])
                                  / float(np.median(builds)) / 1e9 / HBM_PEAK_GBS, 5),
        "finite": bool(np.isfinite(img).all()), "mean_radiance": round(float(img.mean() / total), 5),
        f"rel_mse_vs_{BATCH_REF_SPP}spp_fp64": round(rel, 6),
        "launch": launch or acc.render_launch(),
    }


def build_bytes(n_tris, n_nodes):
    """Algorithmic bytes of one build: per triangle its 12 B of indices and 36 B of vertices
    read and its 36 B reordered triangle written, plus 64 B per wide node written."""
    return n_tris * (12 + 36 + 36) + 64 * n_nodes


def _builder_id(name):
    from alpine_amd import BUILDER_GPU, BUILDER_HOST
    return BUILDER_GPU if name == "gpu" else BUILDER_HOST


def render_pass(which, builder):
    """One timed pass of a render config with stream launches only (no HIP graph, no
    reference render): the workload a per-kernel profile of the path-tracing stages needs."""
    import torch
    from alpine_amd import scenes
    sc = scenes.many_light_scene() if which == "many_light" else scenes.guiding_scene()
    acc = _render_setup(sc, builder)
    out = {"config": which, "frame": [sc.width, sc.height]}
    if which == "many_light":
        # 8 samples in flight per pixel (the bench's autotuning picks 16 for this frame, 5 %
        # faster; the stage profile is taken at 8)
        _timed_render(acc, sc, 8, 0, slots=8)
        dt, img, _ = _timed_render(acc, sc, sc.spp, sc.seed, slots=8)
        out.update(spp=sc.spp, slots=8, seconds=round(dt, 4),
                   samples_per_s=round(sc.width * sc.height * sc.spp / dt, 1))
        return out
    acc.render_reset(sc.width, sc.height, 11, slots=GUIDE_SLOTS)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    acc.render(GUIDE_CHUNK, sc.max_depth)
    torch.cuda.synchronize()
    out["plain_chunk_s"] = round(time.perf_counter() - t0, 4)
    acc.guide_init(seed=1, lr=GUIDE_LR)
    acc.render_reset(sc.width, sc.height, 21, slots=GUIDE_SLOTS)
    t0 = time.perf_counter()
    acc.render(GUIDE_TRAIN_SPP, sc.max_depth, guide=1, guide_warmup=1, train_steps=GUIDE_STEPS,
               guide_frac=GUIDE_FRAC, train_samples=GUIDE_SAMPLES)
    torch.cuda.synchronize()
    out["train_s"] = round(time.perf_counter() - t0, 4)
    t0 = time.perf_counter()
    acc.render(GUIDE_CHUNK, sc.max_depth, guide=2, guide_frac=GUIDE_FRAC)
    torch.cuda.synchronize()
    out["guided_chunk_s"] = round(time.perf_counter() - t0, 4)
    out.update(chunk_spp=GUIDE_CHUNK, train_spp=GUIDE_TRAIN_SPP)
    return out


def forest_build_benchmark(steps=20, warmup=3):
    """Config 4's per-frame rebuild alone: the 32-scene forest (about 12 M triangles), each step
    turning the last 25 % of every scene's triangles on the GPU and rebuilding all 32 trees in
    one GPU forest build (leaf threshold 2), timed on the host clock around synchronised
    steps."""
    import torch
    import alpine_amd as aa
    from alpine_amd import scenes
    B = scenes.BatchedScenes()
    acc = aa.GpuAccelerator(0)
    for v, f in B.meshes:
        acc.append_mesh(v, f)
    acc.update_scene(leaf_threshold=BATCH_LEAF_THRESHOLD, per_shape=True, stats=False)
    rot = B.rotation(2.0 * np.pi / 120.0)
    times = []
    for i in range(warmup + steps):
        torch.cuda.synchronize()
        a = time.perf_counter()
        for k, (first, count) in enumerate(B.anim_range):
            acc.transform_vertices(k, rot, first=first, count=count)
        acc.update_scene(leaf_threshold=BATCH_LEAF_THRESHOLD, per_shape=True, stats=False)
        torch.cuda.synchronize()
        if i >= warmup:
            times.append(time.perf_counter() - a)
    med = float(np.median(times))
    return {"scenes": len(B.meshes), "triangles": B.triangles, "steps": steps,
            "rebuild_ms_median": round(med * 1e3, 2), "rebuild_ms_min": round(min(times) * 1e3, 2),
            "rebuild_mtris_per_s": round(B.triangles / med / 1e6, 1)}


RENDER_CONFIGS = {
    "many_light_pass": lambda a: render_pass("many_light", _builder_id(a.builder)),
    "guided_pass": lambda a: render_pass("guided", _builder_id(a.builder)),
    "cornell": lambda a: cornell_benchmark(_builder_id(a.builder)),
    "many_light": lambda a: many_light_benchmark(_builder_id(a.builder), tune=not a.no_autotune),
    "batched": lambda a: {"batched_32_scenes_512_32spp": batched_benchmark(frames=a.frames,
                                                                          tune=not a.no_autotune)},
    "guided": lambda a: {"guided_vs_plain": guided_benchmark(_builder_id(a.builder))},
    "mlp": lambda a: {"guide_mlp": mlp_benchmark()},
    "forest_build": lambda a: {"forest_rebuild": forest_build_benchmark(a.steps, a.warmup)},
}


def cpu_baseline(verts, tris, org, dir, sample):
    """The reference's binary Bvh (compiled from its sources, oracle/_ref) on every host core
    this process may use (one thread per core, the reference's tile tasks, alpine.cpp:276-285),
    on the whole incoherent ray set; and on one thread, on the first `sample` rays."""
    import oracle
    cores = len(os.sched_getaffinity(0))
    kind = "reference" if oracle.ref_available() else "port"
    t0 = time.perf_counter()
    bvh = oracle.RefBvh(verts, tris) if kind == "reference" else oracle.OracleBvh(verts, tris, 8)
    build_s = time.perf_counter() - t0
    bvh.intersect(org[:4096], dir[:4096], nthreads=cores)  # warm
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        h = bvh.intersect(org, dir, nthreads=cores)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    n1 = min(sample, len(org))
    t0 = time.perf_counter()
    bvh.intersect(org[:n1], dir[:n1], nthreads=1)
    one = time.perf_counter() - t0
    return {
        "value": round(len(org) / med / 1e6, 4), "unit": "Mrays/s", "cores": cores, "kind": kind,
        "sample": f"all {len(org)} incoherent rays, closest-hit, the reference's binary Bvh "
                  f"(9-bin SAH, leaf 8) compiled from its sources, {cores} threads, median of 3 "
                  f"after a warm-up ({med:.2f} s each); build {build_s:.2f} s",
        "one_thread": {"value": round(n1 / one / 1e6, 4), "unit": "Mrays/s", "cores": 1,
                       "sample": f"first {n1} incoherent rays, one pass ({one:.2f} s)"},
    }, h, bvh


def cpu_render_baseline(cores):
    """The reference integrator (alpine.cpp:290-432, restated in the oracle over the reference
    binary Bvh's restatement) on config 1, the Cornell room at 256 x 256 x 16 spp, on `cores`
    threads: the CPU samples/s beside the GPU's."""
    import oracle
    from alpine_amd import scenes
    sc = scenes.cornell_scene()
    ors = oracle.OracleScene(sc.verts, sc.tris, sc.tri_mat, [oracle.make_material(*m) for m in sc.materials],
                             [oracle.make_light(**l) for l in sc.lights], sc.light_sampler,
                             oracle.make_camera(**sc.camera), sc.background)
    ors.render(sc.width, sc.height, 1, sc.max_depth, None, None, sc.rr_depth, cores)  # warm
    states = oracle.sampler_states(sc.width * sc.height, sc.seed)
    t0 = time.perf_counter()
    ors.render(sc.width, sc.height, sc.spp, sc.max_depth, states, None, sc.rr_depth, cores)
    dt = time.perf_counter() - t0
    return {"value": round(sc.width * sc.height * sc.spp / dt, 1), "unit": "samples/s", "cores": cores,
            "kind": "port", "sample": f"Cornell room {sc.width} x {sc.height} x {sc.spp} spp, depth "
                                      f"{sc.max_depth}, {len(sc.tris)} triangles, one pass ({dt:.2f} s)"}


def cpu_baseline_wide(verts, tris, org, dir, sample, cores):
    """The reference's 8-wide WideBvh with the ISPC kernels' arithmetic (restated in the
    oracle; ISPC itself is not available) on this host's cores."""
    import oracle
    t0 = time.perf_counter()
    w = oracle.OracleWideBvh(verts, tris)
    build_s = time.perf_counter() - t0
    n = min(sample, len(org))
    o, d = org[:n], dir[:n]
    w.intersect(o[:4096], d[:4096], nthreads=cores)
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        w.intersect(o, d, nthreads=cores)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    return {"value": round(n / med / 1e6, 4), "unit": "Mrays/s", "cores": cores, "kind": "port",
            "sample": f"first {n} incoherent rays, closest-hit, WideBvh restated (8-wide, leaf 16, "
                      f"ISPC slab/triangle arithmetic, scalar C loops), median of 3; build {build_s:.2f} s"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU sharding is "
                                   "described in DESIGN.md and not launched"}))
        return 2
    import torch
    from alpine_amd import (BUILDER_GPU, BUILDER_HOST, TRI_REFERENCE, TRI_WATERTIGHT, GpuAccelerator,
                            scenes)
    if args.render_config:
        print(json.dumps(RENDER_CONFIGS[args.render_config](args)), flush=True)
        return 0

    log("[bench] generating scene")
    verts, tris = scenes.microbench_scene()
    lo, hi = scenes.scene_bounds(verts)
    org, dir = scenes.incoherent_rays(N_RAYS, lo, hi, seed=8)
    corg, cdir = scenes.coherent_rays(2048, 2048)
    tfar = np.random.default_rng(9).uniform(0, np.linalg.norm(hi - lo), N_RAYS).astype(np.float32)

    log("[bench] building BVH")
    acc = GpuAccelerator(0)
    acc.append_mesh(verts, tris)
    builder = _builder_id(args.builder)
    acc.update_scene(leaf_threshold=LEAF_THRESHOLD, builder=builder)
    st = acc.stats()
    # steady-state rebuild of the same geometry (the first build also loads the kernels)
    rebuild = []
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        acc.update_scene(leaf_threshold=LEAF_THRESHOLD, builder=builder, stats=False)
        torch.cuda.synchronize()
        rebuild.append(time.perf_counter() - t0)
    rebuild_s = float(np.median(rebuild))
    b_build = build_bytes(int(st["triangles"]), int(st["wide_nodes"]))

    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a.T)).cuda()
    o_d, d_d = dev(org), dev(dir)
    co_d, cd_d = dev(corg), dev(cdir)
    tf_d = torch.from_numpy(tfar).cuda()
    out = acc.intersect(o_d, d_d)
    occ = torch.empty(N_RAYS, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()

    tuned = None
    if args.trace_launch:
        from alpine_amd import autotune
        vals = [int(x) for x in args.trace_launch.split(",")]
        tuned = dict(zip(autotune.KNOBS, vals))
        acc.set_trace_config(autotune.make_config(acc.trace_config(), **tuned))
    elif not args.no_autotune:
        # swept on the timed ray set itself (54 shapes x 6 launches of ~2 ms)
        from alpine_amd import autotune
        tuned, _ = autotune.autotune(acc, o_d, d_d, quick=True, reps=5, log=log)
    log("[bench] timing incoherent closest-hit")
    ms, wall_ms = timed_passes(lambda: acc.intersect(o_d, d_d, out=out), args.steps, args.warmup)
    log(f"[bench] incoherent closest {ms:.3f} ms/step")
    ms_coh, _ = timed_passes(lambda: acc.intersect(co_d, cd_d, out=out), args.steps, 2)
    ms_any, _ = timed_passes(lambda: acc.intersect_any(o_d, d_d, out=occ), args.steps, 2)
    ms_any_t, _ = timed_passes(lambda: acc.intersect_any(o_d, d_d, tf_d, out=occ), args.steps, 2)
    acc.set_triangle_test(TRI_WATERTIGHT)
    ms_wt, _ = timed_passes(lambda: acc.intersect(o_d, d_d, out=out), args.steps, 2)
    acc.set_triangle_test(TRI_REFERENCE)

    log("[bench] per-ray host calls (the adapter's Accelerator::intersect path)")
    per_ray = per_ray_host_rate(acc, org, dir)
    log("[bench] instrumented run (byte model)")
    _, n_node, n_tri = acc.intersect_counted(o_d, d_d)
    _, cn_node, cn_tri = acc.intersect_counted(co_d, cd_d)
    _, an_node, an_tri = acc.intersect_counted(o_d, d_d, any_hit=True)
    node_pr, tri_pr = n_node / N_RAYS, n_tri / N_RAYS
    b_coh = B_RAY_IN + B_HIT_OUT + (B_NODE * cn_node + B_TRI * cn_tri) / N_RAYS
    b_any = B_RAY_IN + B_OCC_OUT + (B_NODE * an_node + B_TRI * an_tri) / N_RAYS
    b_closest = B_RAY_IN + B_HIT_OUT + B_NODE * node_pr + B_TRI * tri_pr
    achieved = b_closest * N_RAYS / (ms * 1e-3) / 1e9
    bw_read = hbm_read_gbs()
    bw = hbm_copy_gbs()
    mrays = N_RAYS / (ms * 1e-3) / 1e6
    traffic = _profiled_traffic()

    result = {
        "metric": METRIC,
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "traversal microbench: displaced icosphere (subdiv 8, seed 7) + 200K "
                        "random triangles = 1,510,720 fp32 triangles, 8-wide quantised BVH; "
                        "4,194,304 incoherent rays, closest-hit, tmax = FLT_MAX",
            "triangles": int(st["triangles"]), "rays": N_RAYS, "builder": args.builder,
            "launch": tuned or "default",
        },
        "roofline": {
            # achieved / frac: the algorithmic bytes of the byte model (64 B per node fetch,
            # 36 B per triangle fetch, ray and hit records) per mean kernel time.  The profile
            # shows the kernel limited by VALU issue, and most node fetches hit L2 / the
            # Infinity Cache: the HBM-side bytes it does move are `traffic` (rocprofv3
            # FETCH_SIZE + WRITE_SIZE of the same launch, profiles/), hbm_measured_* below.
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "traffic_source": "profiles/trace_kernel_incoherent_closest_1p5Mtri_4Mrays.json",
            "limiter": "valu issue (rocprofv3 SQ counters, profiles/)",
            "algorithmic_frac": round(achieved / HBM_PEAK_GBS, 4),
            "hbm_measured_gbs": round(traffic / (ms * 1e-3) / 1e9, 1) if traffic else None,
            "hbm_measured_frac": round(traffic / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
            "bytes_per_ray": round(b_closest, 1), "node_fetches_per_ray": round(node_pr, 3),
            "tri_fetches_per_ray": round(tri_pr, 3), "hbm_stream_read_gbs": round(bw_read, 1),
            "frac_of_stream_read": round(achieved / bw_read, 4), "hbm_copy_gbs": round(bw, 1),
        },
        "extra": {
            "wall_ms_per_step": round(wall_ms, 4),
            "coherent_closest_mrays": round(N_RAYS / (ms_coh * 1e-3) / 1e6, 2),
            "coherent_bytes_per_ray": round(b_coh, 1),
            "coherent_frac_hbm": round(b_coh * N_RAYS / (ms_coh * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "incoherent_any_mrays": round(N_RAYS / (ms_any * 1e-3) / 1e6, 2),
            "incoherent_any_tfar_mrays": round(N_RAYS / (ms_any_t * 1e-3) / 1e6, 2),
            "incoherent_closest_watertight_mrays": round(N_RAYS / (ms_wt * 1e-3) / 1e6, 2),
            "any_bytes_per_ray": round(b_any, 1),
            "any_frac_hbm": round(b_any * N_RAYS / (ms_any * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "per_ray_host": per_ray,
            "build_ms_first": round(st["build_ms"], 1), "rebuild_ms": round(rebuild_s * 1e3, 2),
            "rebuild_mtris_per_s": round(st["triangles"] / rebuild_s / 1e6, 1),
            "build_bytes": b_build,
            "rebuild_frac_hbm": round(b_build / rebuild_s / 1e9 / HBM_PEAK_GBS, 5),
            "wide_nodes": int(st["wide_nodes"]),
            "sah_cost": round(st["sah_cost"], 3), "device_mb": round(st["device_bytes"] / 2**20, 1),
        },
        "cpu_baseline": None,
    }
    if not args.no_render:
        log("[bench] render configs")
        result["extra"]["render"] = render_benchmarks(builder, args.frames, tune=not args.no_autotune)
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline")
        cb, href, _ = cpu_baseline(verts, tris, org, dir, args.cpu_sample)
        result["cpu_baseline"] = cb
        result["extra"]["cpu_baseline_wide"] = cpu_baseline_wide(verts, tris, org, dir,
                                                                 args.cpu_sample, cb["cores"])
        result["extra"]["cpu_baseline_render"] = cpu_render_baseline(cb["cores"])
        n = len(href.prim)
        # the last timed pass was watertight; recompute incoherent hits for the parity summary
        acc.intersect(o_d, d_d, out=out)
        torch.cuda.synchronize()
        g = out[0][:n].cpu().numpy().view(np.uint32)
        gt = out[1][:n].cpu().numpy()
        same = (g == href.prim) & (gt == href.t)
        tie = ~same & (gt == href.t)
        result["extra"]["parity_sample"] = {"rays": n, "agree": int(same.sum()),
                                            "exact_t_ties": int(tie.sum()),
                                            "other": int((~same & ~tie).sum())}
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
