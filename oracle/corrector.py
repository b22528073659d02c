"""FP64 numpy restatement of the learned corrector's net -- TEST
INFRASTRUCTURE ONLY (the checker of nmr_hip_corrector_forward / _backward).

Net: conv3d 6->16, ReLU, conv3d 16->16, ReLU, conv3d 16->2; 3x3x3 kernels,
zero padding 1, stride 1.  Parameters in the library's canonical order:
W1[16][6][3][3][3], b1[16], W2[16][16][3][3][3], b2[16], W3[2][16][3][3][3],
b3[2] (the layout of torch.nn.Conv3d weights).  Tensors here are
channels-first (C, Z, Y, X); the library's host arrays are (Z, Y, X, C).

With emulate=True the forward and backward passes round where the GPU
kernels store BF16: input features, weights, the two hidden activations
(after ReLU), the output gradient and the two propagated input gradients
(after the ReLU gate); biases, the output and parameter gradients stay
FP32/FP64.  With emulate=False the passes are plain FP64 mathematics.

Pinned in tests/test_corrector_oracle.py against torch.nn.functional.conv3d
and autograd (FP64, CPU), finite differences, and torch's own float ->
bfloat16 rounding.
"""
from __future__ import annotations

import numpy as np

NPARAMS = 10402
SHAPES = [("W1", (16, 6, 3, 3, 3)), ("b1", (16,)), ("W2", (16, 16, 3, 3, 3)), ("b2", (16,)),
          ("W3", (2, 16, 3, 3, 3)), ("b3", (2,))]


def unpack(params):
    """Canonical vector -> dict of arrays (FP64)."""
    p = np.asarray(params, dtype=np.float64)
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = p[o : o + n].reshape(shp)
        o += n
    assert o == NPARAMS
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SHAPES])


def bf16(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    nan = np.isnan(a)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, (u | 0x400000) & 0xFFFF0000, r)
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def conv3d(x, w, b=None):
    """'same' 3x3x3 convolution (cross-correlation, as conv3d): x (Ci, Z, Y, X),
    w (Co, Ci, 3, 3, 3) -> (Co, Z, Y, X)."""
    ci, nz, ny, nx = x.shape
    xp = np.zeros((ci, nz + 2, ny + 2, nx + 2))
    xp[:, 1:-1, 1:-1, 1:-1] = x
    out = np.zeros((w.shape[0], nz, ny, nx))
    for kz in range(3):
        for ky in range(3):
            for kx in range(3):
                out += np.einsum("oc,czyx->ozyx", w[:, :, kz, ky, kx], xp[:, kz : kz + nz, ky : ky + ny, kx : kx + nx])
    if b is not None:
        out += np.asarray(b)[:, None, None, None]
    return out


def conv3d_input_grad(g, w):
    """Gradient of sum(g * conv3d(x, w)) with respect to x: (Co, ...) -> (Ci, ...)."""
    co, nz, ny, nx = g.shape
    gp = np.zeros((co, nz + 2, ny + 2, nx + 2))
    gp[:, 1:-1, 1:-1, 1:-1] = g
    out = np.zeros((w.shape[1], nz, ny, nx))
    for kz in range(3):
        for ky in range(3):
            for kx in range(3):
                # out[v] += w[:, :, k]^T g[v - (k - 1)]
                sl = gp[:, 2 - kz : 2 - kz + nz, 2 - ky : 2 - ky + ny, 2 - kx : 2 - kx + nx]
                out += np.einsum("oc,ozyx->czyx", w[:, :, kz, ky, kx], sl)
    return out


def conv3d_weight_grad(x, g):
    """Gradient of sum(g * conv3d(x, w)) with respect to w: (Co, Ci, 3, 3, 3)."""
    ci, nz, ny, nx = x.shape
    xp = np.zeros((ci, nz + 2, ny + 2, nx + 2))
    xp[:, 1:-1, 1:-1, 1:-1] = x
    out = np.zeros((g.shape[0], ci, 3, 3, 3))
    for kz in range(3):
        for ky in range(3):
            for kx in range(3):
                out[:, :, kz, ky, kx] = np.einsum("ozyx,czyx->oc", g, xp[:, kz : kz + nz, ky : ky + ny, kx : kx + nx])
    return out


def _r(x, emulate):
    return bf16(x) if emulate else np.asarray(x, dtype=np.float64)


def forward(params, feats, emulate=True, keep=False):
    """feats (Z, Y, X, 6) -> out (Z, Y, X, 2); keep=True also returns the
    stored activations (x, a1, a2) for backward."""
    p = unpack(params)
    w1, w2, w3 = (_r(p[k], emulate) for k in ("W1", "W2", "W3"))
    x = _r(np.moveaxis(np.asarray(feats, dtype=np.float64), -1, 0), emulate)
    a1 = _r(np.maximum(conv3d(x, w1, p["b1"]), 0.0), emulate)
    a2 = _r(np.maximum(conv3d(a1, w2, p["b2"]), 0.0), emulate)
    out = conv3d(a2, w3, p["b3"])
    res = np.moveaxis(out, 0, -1)
    return (res, (x, a1, a2)) if keep else res


def backward(params, feats, grad_out, emulate=True):
    """Gradient of sum(grad_out * forward(feats)) w.r.t. the parameters
    (canonical vector)."""
    p = unpack(params)
    w2, w3 = (_r(p[k], emulate) for k in ("W2", "W3"))
    _, (x, a1, a2) = forward(params, feats, emulate, keep=True)
    g3 = _r(np.moveaxis(np.asarray(grad_out, dtype=np.float64), -1, 0), emulate)
    g2 = _r(conv3d_input_grad(g3, w3) * (a2 > 0), emulate)
    g1 = _r(conv3d_input_grad(g2, w2) * (a1 > 0), emulate)
    d = {
        "W3": conv3d_weight_grad(a2, g3), "b3": g3.sum(axis=(1, 2, 3)),
        "W2": conv3d_weight_grad(a1, g2), "b2": g2.sum(axis=(1, 2, 3)),
        "W1": conv3d_weight_grad(x, g1), "b1": g1.sum(axis=(1, 2, 3)),
    }
    return pack(d)


def random_params(seed, scale=1.0):
    """He-normal weights (as the library initialises them, any generator)
    and small random biases, for tests."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shp in SHAPES:
        if name.startswith("W"):
            fan = shp[1] * 27
            d[name] = rng.normal(0.0, scale * np.sqrt(2.0 / fan), shp)
        else:
            d[name] = rng.normal(0.0, 0.05, shp)
    return pack(d)
