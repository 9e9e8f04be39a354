"""Guiding MLP on the GPU (BF16 MFMA forward / backward / Adam) against the FP64 oracle."""
import numpy as np
import pytest

import alpine_amd as aa
from oracle import guide as G

pytestmark = pytest.mark.gpu


def _acc():
    acc = aa.GpuAccelerator(0)
    acc.guide_init(seed=3, lr=1e-3)
    return acc


def _data(m, seed=0):
    rng = np.random.default_rng(seed)
    p01 = rng.random((m, 3))
    n = rng.standard_normal((m, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    wo = rng.standard_normal((m, 3))
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    feat = G.encode(p01, n, wo).astype(np.float32)
    d = rng.standard_normal((m, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    w = rng.exponential(1.0, m).astype(np.float32)
    return feat, d.astype(np.float32), w


def _cuda(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_forward_matches_oracle():
    print("[gpu] guide forward", flush=True)
    acc = _acc()
    params = acc.guide_params().astype(np.float64)
    feat, _, _ = _data(1000)
    out = acc.guide_forward(_cuda(feat)).cpu().numpy()
    ref, _ = G.forward(params, feat.astype(np.float64), bf=True)
    err = np.abs(out[:, :40] - ref[:, :40]).max() / np.abs(ref[:, :40]).max()
    print("forward max rel err", err)
    assert err < 2e-2


def test_fused_inference_matches_oracle():
    """Inference with the encoding computed in registers (the renderer's path) vs the oracle."""
    print("[gpu] guide fused inference", flush=True)
    acc = _acc()
    params = acc.guide_params().astype(np.float64)
    rng = np.random.default_rng(5)
    m = 1000  # not a multiple of the 16-query tile
    p01 = rng.random((m, 3))
    n = rng.standard_normal((m, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    wo = rng.standard_normal((m, 3))
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    q = np.concatenate([p01, n, wo], 1).astype(np.float32)
    out = acc.guide_infer(_cuda(q)).cpu().numpy()
    ref, _ = G.forward(params, G.encode(q[:, 0:3].astype(np.float64), q[:, 3:6], q[:, 6:9]), bf=True)
    err = np.abs(out[:, :40] - ref[:, :40]).max() / np.abs(ref[:, :40]).max()
    print("fused inference max rel err", err)
    assert err < 2e-2
    # same network as the unfused forward pass on the oracle's features
    feat = G.encode(q[:, 0:3].astype(np.float64), q[:, 3:6], q[:, 6:9]).astype(np.float32)
    out2 = acc.guide_forward(_cuda(feat)).cpu().numpy()
    assert np.abs(out - out2)[:, :40].max() / np.abs(out2[:, :40]).max() < 2e-2


def test_gradient_matches_oracle():
    print("[gpu] guide gradient", flush=True)
    acc = _acc()
    params = acc.guide_params().astype(np.float64)
    feat, d, w = _data(3000, seed=1)
    loss, g = acc.guide_train_step(_cuda(feat), _cuda(d), _cuda(w), apply=False, want_grad=True)
    gref = G.grad(params, feat.astype(np.float64), d.astype(np.float64), w.astype(np.float64), bf=True)
    lref = G.loss(params, feat.astype(np.float64), d.astype(np.float64), w.astype(np.float64), bf=True)
    cos = float(g @ gref / (np.linalg.norm(g) * np.linalg.norm(gref)))
    rel = float(np.linalg.norm(g - gref) / np.linalg.norm(gref))
    print("grad cos", cos, "rel l2", rel, "loss", loss, lref)
    assert cos > 0.999 and rel < 3e-2
    assert loss == pytest.approx(lref, rel=1e-2)
    for l in range(5):  # every layer's block agrees
        a, b = g[G.W_OFF[l]:G.W_OFF[l + 1]], gref[G.W_OFF[l]:G.W_OFF[l + 1]]
        assert float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b))) > 0.995, l


def test_fused_training_step_matches_oracle():
    """The renderer's training path (encoding, forward, head and backward-data in one kernel)
    against the FP64 oracle with the same BF16 rounding points."""
    print("[gpu] guide fused training gradient", flush=True)
    acc = _acc()
    params = acc.guide_params().astype(np.float64)
    rng = np.random.default_rng(9)
    m = 2999  # not a multiple of the 16-query tile
    p01 = rng.random((m, 3))
    n = rng.standard_normal((m, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    wo = rng.standard_normal((m, 3))
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    q = np.concatenate([p01, n, wo], 1).astype(np.float32)
    d = rng.standard_normal((m, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    w = rng.exponential(1.0, m).astype(np.float32)
    loss, g = acc.guide_train_queries(_cuda(q), _cuda(d), _cuda(w), apply=False, want_grad=True)
    feat = G.encode(q[:, 0:3].astype(np.float64), q[:, 3:6], q[:, 6:9])
    gref = G.grad(params, feat, d.astype(np.float64), w.astype(np.float64), bf=True)
    lref = G.loss(params, feat, d.astype(np.float64), w.astype(np.float64), bf=True)
    cos = float(g @ gref / (np.linalg.norm(g) * np.linalg.norm(gref)))
    rel = float(np.linalg.norm(g - gref) / np.linalg.norm(gref))
    print("fused grad cos", cos, "rel l2", rel, "loss", loss, lref)
    assert cos > 0.999 and rel < 3e-2
    assert loss == pytest.approx(lref, rel=1e-2)


def test_adam_step_matches_oracle():
    print("[gpu] guide adam", flush=True)
    acc = _acc()
    p0 = acc.guide_params().astype(np.float64)
    feat, d, w = _data(2000, seed=2)
    _, g = acc.guide_train_step(_cuda(feat), _cuda(d), _cuda(w), apply=True, want_grad=True)
    p1 = acc.guide_params().astype(np.float64)
    pref, _, _ = G.adam(p0, g.astype(np.float64), np.zeros_like(p0), np.zeros_like(p0), 1)
    np.testing.assert_allclose(p1, pref, rtol=0, atol=1e-6)


def test_training_fits_a_directional_target():
    """Samples concentrated around a direction that depends on position: the trained
    mixture assigns them much higher likelihood than the initial one."""
    print("[gpu] guide training", flush=True)
    acc = _acc()
    rng = np.random.default_rng(4)
    m = 8192
    p01 = rng.random((m, 3))
    n = np.tile([0.0, 0.0, 1.0], (m, 1))
    wo = np.tile([0.0, 0.0, 1.0], (m, 1))
    feat = G.encode(p01, n, wo).astype(np.float32)
    tgt = np.stack([np.cos(2 * np.pi * p01[:, 0]), np.sin(2 * np.pi * p01[:, 0]),
                    0.3 * np.ones(m)], 1)
    tgt /= np.linalg.norm(tgt, axis=1, keepdims=True)
    d = tgt + 0.15 * rng.standard_normal((m, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    w = np.ones(m, np.float32)
    f_d, d_d, w_d = _cuda(feat), _cuda(d), _cuda(w)
    first, _ = acc.guide_train_step(f_d, d_d, w_d, apply=False)
    for _ in range(300):
        last, _ = acc.guide_train_step(f_d, d_d, w_d, apply=True)
    print("nll before", first, "after", last)
    assert last < first - 2.0
