"""CPU restatement of the learned pair ranker's network -- TEST INFRASTRUCTURE.

Imported only by tests/ (never by the product).  The network is
24 -> 128 -> 128 -> 1 with ReLU; the parameter vector layout, the feature
normalisation and the target transform follow beam_amd/csrc/ranker.hpp.

Precision semantics of the HIP path, restated exactly:
  * forward: normalised features, W1 and W2 are rounded to BF16 (round to
    nearest even), the first hidden layer's post-ReLU activations are rounded
    to BF16 before the second product, accumulation is FP32 (here FP64, which
    only tightens the comparison);
  * backward: FP32 from the stored FP32 activations and the FP32 master
    weights (the usual mixed-precision gradient of the BF16 forward);
  * Adam: beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected.
"""
from __future__ import annotations

import numpy as np

FEAT, HID = 24, 128
OFF_W1 = 0
OFF_B1 = OFF_W1 + FEAT * HID
OFF_W2 = OFF_B1 + HID
OFF_B2 = OFF_W2 + HID * HID
OFF_W3 = OFF_B2 + HID
OFF_B3 = OFF_W3 + HID
PARAMS = OFF_B3 + 1
NORM = 2 * FEAT + 2


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to BF16 (nearest even), returned as float32."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (r.astype(np.uint32) << 16).view(np.float32)


def unpack(params: np.ndarray):
    p = np.asarray(params)
    return (p[OFF_W1:OFF_B1].reshape(FEAT, HID), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(HID, HID),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], p[OFF_B3])


def normalise(norm: np.ndarray, feats: np.ndarray) -> np.ndarray:
    norm = np.asarray(norm, dtype=np.float32)
    return ((np.asarray(feats, dtype=np.float32) - norm[:FEAT]) * norm[FEAT:2 * FEAT]).astype(np.float32)


def forward(params, norm, feats, mixed=True):
    """Returns (y, X, H1, H2); mixed=False computes everything in FP64."""
    w1, b1, w2, b2, w3, b3 = unpack(params)
    if mixed:
        x = normalise(norm, feats)
        xq, w1q, w2q = bf16(x), bf16(w1.astype(np.float32)), bf16(w2.astype(np.float32))
        b1, b2, w3, b3 = (np.asarray(b1, np.float32), np.asarray(b2, np.float32), np.asarray(w3, np.float32),
                          np.float32(b3))
    else:
        norm64 = np.asarray(norm, dtype=np.float64)
        x = (np.asarray(feats, dtype=np.float64) - norm64[:FEAT]) * norm64[FEAT:2 * FEAT]
        xq, w1q, w2q = x, w1, w2
    h1 = np.maximum(xq.astype(np.float64) @ w1q.astype(np.float64) + b1, 0.0)
    h1q = bf16(h1.astype(np.float32)).astype(np.float64) if mixed else h1
    h2 = np.maximum(h1q @ w2q.astype(np.float64) + b2, 0.0)
    y = h2 @ w3.astype(np.float64) + np.float64(b3)
    return y, x.astype(np.float64), h1, h2


def standardised_target(norm, target) -> np.ndarray:
    norm = np.asarray(norm, dtype=np.float64)
    return (np.asarray(target, dtype=np.float64) - norm[2 * FEAT]) * norm[2 * FEAT + 1]


def loss(params, norm, feats, target, mixed=True) -> float:
    y, _, _, _ = forward(params, norm, feats, mixed)
    d = y - standardised_target(norm, target)
    return float(np.mean(d * d))


def gradient(params, norm, feats, target, mixed=True):
    """Gradient of the mean squared error over the batch; returns (grad, loss)."""
    _, _, w2, _, w3, _ = unpack(params)
    y, x, h1, h2 = forward(params, norm, feats, mixed)
    n = len(y)
    d = y - standardised_target(norm, target)
    dy = 2.0 * d / n
    g = np.zeros(PARAMS, dtype=np.float64)
    g[OFF_W3:OFF_B3] = h2.T @ dy
    g[OFF_B3] = dy.sum()
    dh2 = dy[:, None] * w3.astype(np.float64)[None, :] * (h2 > 0)
    g[OFF_W2:OFF_B2] = (h1.T @ dh2).ravel()
    g[OFF_B2:OFF_W3] = dh2.sum(axis=0)
    dh1 = (dh2 @ w2.astype(np.float64).T) * (h1 > 0)
    g[OFF_W1:OFF_B1] = (x.T @ dh1).ravel()
    g[OFF_B1:OFF_W2] = dh1.sum(axis=0)
    return g, float(np.mean(d * d))


def adam(params, grad, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (step counts from 1); returns (params, m, v) as float32."""
    g = np.asarray(grad, dtype=np.float32)
    m = (np.float32(b1) * m + np.float32(1 - b1) * g).astype(np.float32)
    v = (np.float32(b2) * v + np.float32(1 - b2) * g * g).astype(np.float32)
    bc1 = np.float32(1 - b1 ** step)
    bc2 = np.float32(1 - b2 ** step)
    p = (params - np.float32(lr) * (m / bc1) / (np.sqrt(v / bc2) + np.float32(eps))).astype(np.float32)
    return p, m, v


def init_params(seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    p = np.zeros(PARAMS, dtype=np.float32)
    for off, fan_in, count in ((OFF_W1, FEAT, FEAT * HID), (OFF_W2, HID, HID * HID), (OFF_W3, HID, HID)):
        lim = np.sqrt(6.0 / fan_in)
        p[off:off + count] = rng.uniform(-lim, lim, count).astype(np.float32)
    p[OFF_B1:OFF_W2] = rng.uniform(-0.1, 0.1, HID).astype(np.float32)
    p[OFF_B2:OFF_W3] = rng.uniform(-0.1, 0.1, HID).astype(np.float32)
    p[OFF_B3] = np.float32(0.05)
    return p
