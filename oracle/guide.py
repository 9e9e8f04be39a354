"""ORACLE -- test infrastructure only.

FP64 numpy restatement of the path-guiding network (alpine_amd/csrc/guide_core.h,
guide.hip): 96 -> 64 -> 64 -> 64 -> 64 -> 48, ReLU, 8-lobe von Mises-Fisher
mixture head, weighted negative log-likelihood loss, backward pass and Adam.

``bf16=True`` rounds at the points where the GPU kernels store BF16 (features,
weights, hidden activations, output gradients, hidden gradients); everything
else is FP64.  ``bf16=False`` is the exact FP64 network, which the tests pin
with central finite differences.  No reference counterpart exists (the
reference has no guiding); this file is the whole specification.
"""
from __future__ import annotations

import numpy as np

IN, HID, OUT, LOBES, FREQS = 96, 64, 48, 8, 12
K = [IN, HID, HID, HID, HID]
N = [HID, HID, HID, HID, OUT]
W_OFF = [0, 6144, 10240, 14336, 18432, 21504]
B_OFF = [21504, 21568, 21632, 21696, 21760, 21808]
PARAMS = 21808
LOGK_MIN, LOGK_MAX = -3.0, 7.0


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(params):
    Ws = [params[W_OFF[l]:W_OFF[l] + K[l] * N[l]].reshape(N[l], K[l]) for l in range(5)]
    bs = [params[B_OFF[l]:B_OFF[l] + N[l]] for l in range(5)]
    return Ws, bs


def encode(p01, n, wo):
    """Features (guide_encode): p, sin/cos(2^k pi p), n, wo, zero padding."""
    m = len(p01)
    f = np.zeros((m, IN), np.float64)
    f[:, 0:3] = p01
    for a in range(3):
        for k in range(FREQS):
            x = (2.0 ** k) * np.pi * p01[:, a]
            f[:, 3 + 24 * a + 2 * k] = np.sin(x)
            f[:, 3 + 24 * a + 2 * k + 1] = np.cos(x)
    f[:, 75:78] = n
    f[:, 78:81] = wo
    return f


def forward(params, feat, bf=True):
    Ws, bs = unpack(np.asarray(params, np.float64))
    r = bf16 if bf else (lambda x: np.asarray(x, np.float64))
    a = r(feat)
    acts = [a]
    for l in range(5):
        z = a @ r(Ws[l]).T + bs[l]
        if l < 4:
            a = r(np.maximum(z, 0.0))
            acts.append(a)
        else:
            out = z
    return out, acts


def decode(o):
    a = o[:, 0:40:5]
    a = a - a.max(axis=1, keepdims=True)
    pi = np.exp(a)
    pi /= pi.sum(axis=1, keepdims=True)
    kappa = np.exp(np.clip(o[:, 1:40:5], LOGK_MIN, LOGK_MAX))
    v = np.stack([o[:, 2:40:5], o[:, 3:40:5], o[:, 4:40:5]], -1)
    l = np.linalg.norm(v, axis=-1, keepdims=True)
    mu = np.where(l > 1e-6, v / np.maximum(l, 1e-30), np.array([0.0, 0.0, 1.0]))
    return pi, kappa, mu, v, l


def vmf_log(kappa, c):
    return np.log(kappa) - np.log(2 * np.pi) - np.log1p(-np.exp(-2 * kappa)) + kappa * (c - 1)


def nll(o, w):
    """-log q(w | o) per sample."""
    pi, kappa, mu, _, _ = decode(o)
    c = np.einsum("mij,mj->mi", mu, w)
    lv = np.log(np.maximum(pi, 1e-30)) + vmf_log(kappa, c)
    mx = lv.max(axis=1, keepdims=True)
    return -(mx[:, 0] + np.log(np.exp(lv - mx).sum(axis=1)))


def nll_grad(o, w, scale):
    """d(scale * -log q)/d o (guide_nll_grad)."""
    pi, kappa, mu, v, l = decode(o)
    c = np.einsum("mij,mj->mi", mu, w)
    lv = np.log(np.maximum(pi, 1e-30)) + vmf_log(kappa, c)
    mx = lv.max(axis=1, keepdims=True)
    logq = mx + np.log(np.exp(lv - mx).sum(axis=1, keepdims=True))
    r = np.exp(lv - logq)
    g = np.zeros_like(o)
    s = scale[:, None]
    g[:, 0:40:5] = s * (pi - r)
    e2 = np.exp(-2 * kappa)
    dlk = 1.0 / kappa - 2 * e2 / (1 - e2) + (c - 1)
    ok = o[:, 1:40:5]
    inside = (ok > LOGK_MIN) & (ok < LOGK_MAX)
    g[:, 1:40:5] = np.where(inside, s * (-r * dlk * kappa), 0.0)
    dmu = w[:, None, :] * (-r * kappa)[:, :, None]
    dv = (dmu - mu * np.sum(mu * dmu, -1, keepdims=True)) / np.maximum(l, 1e-30)
    dv = np.where(l > 1e-6, dv, 0.0)
    for j in range(3):
        g[:, 2 + j:40:5] = s * dv[:, :, j]
    return g


def loss(params, feat, dirs, weight, bf=False):
    out, _ = forward(params, feat, bf)
    wsum = weight.sum()
    return float(np.sum(weight / wsum * nll(out, dirs)))


def grad(params, feat, dirs, weight, bf=True):
    """Gradient of loss() (with the GPU's BF16 rounding points when bf=True)."""
    params = np.asarray(params, np.float64)
    r = bf16 if bf else (lambda x: np.asarray(x, np.float64))
    Ws, _ = unpack(params)
    out, acts = forward(params, feat, bf)
    wsum = weight.sum()
    dz = r(nll_grad(out, dirs, weight / wsum))
    g = np.zeros(PARAMS)
    for l in range(4, -1, -1):
        g[W_OFF[l]:W_OFF[l] + K[l] * N[l]] = (dz.T @ acts[l]).reshape(-1)
        g[B_OFF[l]:B_OFF[l] + N[l]] = dz.sum(axis=0)
        if l > 0:
            da = dz @ r(Ws[l])
            dz = r(np.where(acts[l] > 0, da, 0.0))
    return g


def adam(params, g, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = params - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
    return p, m, v
