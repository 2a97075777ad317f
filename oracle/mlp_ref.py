"""CPU oracle of the learned presizing (test infrastructure only).

Restates, in numpy float64, what wtet_amd's presizing kernels compute:
per-cloud features (512-bin occupancy histogram of the top 9 bits of the
30-bit Morton code with the 10-bit floor quantization of
src/morton.rs:22-35, mean and covariance of the box-normalized
coordinates, log n) and the 522 -> 128 -> 128 -> 3 ReLU MLP with its MSE
loss and gradients, in the parameter layout of include/wtet.h.

The gradient is pinned by central finite differences in the tests (a
mathematical check, independent of this file's formulas)."""
from __future__ import annotations

import numpy as np

F, H, O = 522, 128, 3
OFF = {}
_o = 0
for name, size in [("fmean", F), ("fstd", F), ("tmean", O), ("tstd", O), ("W1", F * H), ("b1", H),
                   ("W2", H * H), ("b2", H), ("W3", H * O), ("b3", O), ("margin", O)]:
    OFF[name] = (_o, _o + size)
    _o += size
NPARAMS = _o
TRAIN = (OFF["W1"][0], OFF["b3"][1])


def part(p, name):
    a, b = OFF[name]
    v = p[a:b]
    if name == "W1":
        return v.reshape(F, H)
    if name == "W2":
        return v.reshape(H, H)
    if name == "W3":
        return v.reshape(H, O)
    return v


def _expand10(v):
    v = v.astype(np.uint64) & 0x3FF
    v = (v | (v << 16)) & 0x030000FF
    v = (v | (v << 8)) & 0x0300F00F
    v = (v | (v << 4)) & 0x030C30C3
    v = (v | (v << 2)) & 0x09249249
    return v


def features(points: np.ndarray) -> np.ndarray:
    """Features of one cloud (raw f32 points), as the device computes them
    (f32 box normalization and quantization, f64 moments)."""
    p = np.asarray(points, np.float32)
    lo, hi = p.min(0), p.max(0)
    ext = (hi - lo).astype(np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        u = np.where(ext > 0, (p - lo) / np.where(ext > 0, ext, np.float32(1)), np.float32(0))
    u = np.clip(u.astype(np.float32), np.float32(0), np.float32(1))
    q = (u * np.float32(1023.0)).astype(np.uint32)
    code = (_expand10(q[:, 2]) << 2) | (_expand10(q[:, 1]) << 1) | _expand10(q[:, 0])
    n = len(p)
    hist = np.bincount((code >> 21).astype(np.int64), minlength=512).astype(np.float64) / n
    ud = u.astype(np.float64)
    mean = ud.mean(0)
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    cov = [np.mean(ud[:, a] * ud[:, b]) - mean[a] * mean[b] for a, b in pairs]
    return np.concatenate([hist, mean, cov, [np.log(n)]])


def forward(p, X):
    """Standardized outputs and the activations, from raw features X."""
    p = np.asarray(p, np.float64)
    x = (X - part(p, "fmean")) / part(p, "fstd")
    z1 = x @ part(p, "W1") + part(p, "b1")
    h1 = np.maximum(z1, 0)
    z2 = h1 @ part(p, "W2") + part(p, "b2")
    h2 = np.maximum(z2, 0)
    out = h2 @ part(p, "W3") + part(p, "b3")
    return out, (x, h1, h2)


def predict(p, X):
    """Un-standardized log-space predictions plus the margin."""
    out, _ = forward(p, X)
    return out * part(p, "tstd") + part(p, "tmean") + part(p, "margin")


def loss_grad(p, X, Y):
    """MSE over standardized targets and its gradient (trainable range)."""
    p = np.asarray(p, np.float64)
    out, (x, h1, h2) = forward(p, X)
    y = (Y - part(p, "tmean")) / part(p, "tstd")
    e = out - y
    loss = np.mean(e * e)
    d3 = 2 * e / e.size
    g = np.zeros(NPARAMS)
    g[OFF["W3"][0]:OFF["W3"][1]] = (h2.T @ d3).ravel()
    g[OFF["b3"][0]:OFF["b3"][1]] = d3.sum(0)
    d2 = (d3 @ part(p, "W3").T) * (h2 > 0)
    g[OFF["W2"][0]:OFF["W2"][1]] = (h1.T @ d2).ravel()
    g[OFF["b2"][0]:OFF["b2"][1]] = d2.sum(0)
    d1 = (d2 @ part(p, "W2").T) * (h1 > 0)
    g[OFF["W1"][0]:OFF["W1"][1]] = (x.T @ d1).ravel()
    g[OFF["b1"][0]:OFF["b1"][1]] = d1.sum(0)
    return loss, g
