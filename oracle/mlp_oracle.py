"""CPU oracle of the learned surrogate network (TEST INFRASTRUCTURE: imported only by tests/,
__graft_entry__.smoke() and bench.py's CPU baseline; never by the product path).

The network has no counterpart in the reference; this restates, in float64 numpy, what the
library computes:
  features   per vertex and strut-neighbour slot (nbr, rest L): strain_scale * (|d| - L) / L *
             d / |d| and (v_nbr - v) * H / L with d = x_nbr - x (empty slot: zeros)
  forward    z1 = W1 x + b1, h1 = gelu(z1), ... h3 = gelu(z3), y = W4 h3 + b4
             gelu(z) = z clamp(0.5 + z (a - b min(|z|, z*)), 0, 1), a = 1/z*, b = 1/(2 z*^2)
             (the library's C1 piecewise-quadratic GELU, within 0.0197 of z Phi(z))
  loss       mean over rows and the 3 outputs of (y - t)^2
  backward   the analytic gradient of that loss
With emulate=True the rounding points of the GPU path are applied: weights, inputs, hidden
activations, dY and dZ rounded to bf16 (nearest even), so the remaining difference to the GPU is
only f32-versus-f64 accumulation.

Pinned by tests/test_surrogate_oracle.py: the analytic gradient against central finite
differences (float64, emulate=False), the activation against the exact GELU z Phi(z) and its
saturation, hand-computed features.
"""
from __future__ import annotations

import numpy as np

HID = 128
FEAT = 72
OUT = 3
OFF_W1 = 0
OFF_B1 = OFF_W1 + HID * FEAT
OFF_W2 = OFF_B1 + HID
OFF_B2 = OFF_W2 + HID * HID
OFF_W3 = OFF_B2 + HID
OFF_B3 = OFF_W3 + HID * HID
OFF_W4 = OFF_B3 + HID
OFF_B4 = OFF_W4 + OUT * HID
N_PARAMS = OFF_B4 + OUT

# the kernels' f32 constants: 1/z*, a and b z*, with min(|z|, z*) formed as z* clamp(|z| / z*, 0, 1)
XS = 2.4629791631638027
IXS = float(np.float32(1.0 / XS))
A = float(np.float32(0.4060123670374283))
BX = float(np.float32(0.08242302109366773 * XS))
B = float(np.float32(0.08242302109366773))


def bf16(a):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _s_unclamped(z):
    t = np.clip(np.abs(z) * IXS, 0.0, 1.0)
    return 0.5 + z * (A - BX * t)


def gelu(z):
    z = np.asarray(z, np.float64)
    return z * np.clip(_s_unclamped(z), 0.0, 1.0)


def gelu_grad(z):
    z = np.asarray(z, np.float64)
    ds = np.where(np.abs(z) < XS, A - 2.0 * B * np.abs(z), 0.0)
    return np.clip(_s_unclamped(z), 0.0, 1.0) + z * ds


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return [(p[OFF_W1:OFF_B1].reshape(HID, FEAT), p[OFF_B1:OFF_W2]),
            (p[OFF_W2:OFF_B2].reshape(HID, HID), p[OFF_B2:OFF_W3]),
            (p[OFF_W3:OFF_B3].reshape(HID, HID), p[OFF_B3:OFF_W4]),
            (p[OFF_W4:OFF_B4].reshape(OUT, HID), p[OFF_B4:N_PARAMS])]


def features(pos, vel, nbr, rest, H, strain_scale):
    """(n, 72) features of vertices with positions/velocities (n, 3) and slot tables nbr (n, 12;
    -1 empty) and rest (n, 12)."""
    pos = np.asarray(pos, np.float64)
    vel = np.asarray(vel, np.float64)
    n = pos.shape[0]
    out = np.zeros((n, 12, 6))
    m = nbr >= 0
    j = np.where(m, nbr, 0)
    d = pos[j] - pos[:, None, :]
    ln = np.linalg.norm(d, axis=2)
    L = np.where(m, rest, 1.0).astype(np.float64)
    st = np.where(ln > 0, strain_scale * ((ln - L) / L) / np.where(ln > 0, ln, 1.0), 0.0)
    out[..., 0:3] = st[..., None] * d
    out[..., 3:6] = (vel[j] - vel[:, None, :]) * (H / L)[..., None]
    out[~m] = 0.0
    return out.reshape(n, 72)


def forward(p, X, emulate=False):
    """Outputs (rows, 3) and the cache for backward."""
    layers = unpack(p)
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    h = r(X)
    cache = {"h": [h], "z": []}
    for li, (W, b) in enumerate(layers):
        z = h @ r(W).T + b
        if li < 3:
            cache["z"].append(z)
            h = r(gelu(z))
            cache["h"].append(h)
        else:
            y = z
    return y, cache


def loss_and_grad(p, X, T, emulate=False):
    """(loss, grad (N_PARAMS,), Y)."""
    layers = unpack(p)
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    Y, cache = forward(p, X, emulate)
    T = np.asarray(T, np.float64)
    rows = Y.shape[0]
    loss = float(np.mean((Y - T) ** 2))
    g = np.zeros(N_PARAMS)
    dz = r(2.0 * (Y - T) / (3.0 * rows))
    offs = [(OFF_W1, OFF_B1), (OFF_W2, OFF_B2), (OFF_W3, OFF_B3), (OFF_W4, OFF_B4)]
    for li in (3, 2, 1, 0):
        W, _ = layers[li]
        hprev = cache["h"][li]
        ow, ob = offs[li]
        g[ow:ow + W.size] = (dz.T @ hprev).reshape(-1)
        g[ob:ob + W.shape[0]] = dz.sum(axis=0)
        if li > 0:
            dh = dz @ r(W)
            dz = r(dh * gelu_grad(cache["z"][li - 1]))
    return loss, g, Y


def init_params(seed=0):
    """Uniform +-1/sqrt(fan_in) weights and biases (the library's initialisation law, not its stream)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS)
    for (ow, ob, fan_in, rows) in ((OFF_W1, OFF_B1, FEAT, HID), (OFF_W2, OFF_B2, HID, HID), (OFF_W3, OFF_B3, HID, HID),
                                   (OFF_W4, OFF_B4, HID, OUT)):
        b = 1.0 / np.sqrt(fan_in)
        p[ow:ob] = rng.uniform(-b, b, size=ob - ow)
        p[ob:ob + rows] = rng.uniform(-b, b, size=rows)
    return p
