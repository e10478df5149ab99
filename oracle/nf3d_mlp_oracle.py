"""FP64 restatement of the learned-PSD network (test infrastructure only).

This module is the checker for the BF16 matrix-core kernels of
neo_fft3d_amd/csrc/nf3d_mlp.hip: only the tests import it.  The network has no
counterpart in the reference (it replaces the single noise pattern of
src/helper.cpp:57-132 with a per-block one), so its definition is fixed here:

    h1 = gelu(x W1 + b1),  h2 = gelu(h1 W2 + b2),  y = h2 W3 + b3
    gelu(z) = z Phi(z)  (exact, erf form)
    loss = mean over rows and outputs of  y + t exp(-y)      (Whittle)

y is the predicted log noise PSD of a block, t the periodogram of the noise
alone.  For each output, E[t] = S minimises E[y + t e^-y] at y = ln S.

Pinning: the analytic gradient is checked against central finite differences
(``fd_check``), the forward against a direct evaluation with
scipy.special.erf, and the loss minimiser against the closed form
y* = ln(mean t) (``whittle_minimiser``).
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

SQRT1_2 = 0.7071067811865476
INV_SQRT_2PI = 0.3989422804014327


def bf16_round(a: np.ndarray) -> np.ndarray:
    """Round float values to BF16 (nearest, ties to even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z * SQRT1_2))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z * SQRT1_2)) + z * INV_SQRT_2PI * np.exp(-0.5 * z * z)


def unpack(flat: np.ndarray, n_in: int, nh: int, n_out: int):
    """W1 [n_in][nh], b1, W2 [nh][nh], b2, W3 [nh][n_out], b3 from the flat host order."""
    sizes = [n_in * nh, nh, nh * nh, nh, nh * n_out, n_out]
    shapes = [(n_in, nh), (nh,), (nh, nh), (nh,), (nh, n_out), (n_out,)]
    out, at = [], 0
    for s, sh in zip(sizes, shapes):
        out.append(np.asarray(flat[at:at + s], dtype=np.float64).reshape(sh))
        at += s
    return out


def pack(parts) -> np.ndarray:
    return np.concatenate([np.asarray(p, dtype=np.float64).ravel() for p in parts])


def forward(params, x, bf16: bool = False):
    """Activations of the network; with bf16, the operands the kernels round
    (weights, inputs of each product) are rounded to BF16."""
    W1, b1, W2, b2, W3, b3 = params
    r = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    x = r(x)
    z1 = x @ r(W1) + b1
    h1 = gelu(z1)
    z2 = r(h1) @ r(W2) + b2
    h2 = gelu(z2)
    y = r(h2) @ r(W3) + b3
    return dict(x=x, z1=z1, h1=h1, z2=z2, h2=h2, y=y)


def whittle_loss(y, t) -> float:
    return float(np.mean(y + t * np.exp(-y)))


def loss_of(params, x, t) -> float:
    return whittle_loss(forward(params, x)["y"], t)


def backward(params, x, t, bf16: bool = False):
    """Gradients of the Whittle loss, in the flat host order."""
    W1, b1, W2, b2, W3, b3 = params
    r = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    a = forward(params, x, bf16)
    n = a["y"].size
    dy = (1.0 - t * np.exp(-a["y"])) / n
    gW3 = r(a["h2"]).T @ r(dy)
    gb3 = dy.sum(0)
    dz2 = (r(dy) @ r(W3).T) * gelu_grad(a["z2"])
    gW2 = r(a["h1"]).T @ r(dz2)
    gb2 = dz2.sum(0)
    dz1 = (r(dz2) @ r(W2).T) * gelu_grad(a["z1"])
    gW1 = a["x"].T @ r(dz1)
    gb1 = dz1.sum(0)
    return pack([gW1, gb1, gW2, gb2, gW3, gb3])


def fd_check(params, x, t, idx, h: float = 1e-6):
    """Central differences of the loss at flat parameter indices `idx`."""
    flat = pack(params)
    shapes = [p.shape for p in params]
    out = []
    for i in idx:
        vals = []
        for s in (+h, -h):
            f = flat.copy()
            f[i] += s
            parts, at = [], 0
            for sh in shapes:
                k = int(np.prod(sh))
                parts.append(f[at:at + k].reshape(sh))
                at += k
            vals.append(loss_of(parts, x, t))
        out.append((vals[0] - vals[1]) / (2 * h))
    return np.array(out)


def whittle_minimiser(t: np.ndarray) -> np.ndarray:
    """argmin_y mean_rows(y + t e^-y) per output: ln of the mean of t."""
    return np.log(np.mean(t, axis=0))

