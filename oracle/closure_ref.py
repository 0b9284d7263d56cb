"""FP64 numpy oracle of the learned closure network (test infrastructure only).

The closure has no counterpart in the reference; this file states the
network the HIP kernels implement (lbman3d_amd/csrc/closure.hip) so that the
kernels can be checked against it: three periodic 3x3x3 convolutions
8 -> H -> H -> 5 with ReLU after the first two, the mean-squared loss over
V cells x 5 channels, and its exact gradient. The gradient is pinned by
central finite differences in tests/test_closure_oracle.py; the forward pass
by a brute-force per-cell loop on tiny grids.

Arrays are grid-shaped (nz, ny, nx, C). Parameters are one flat vector,
layer by layer W[27][cin][cout] (stencil offset (dz+1)*9 + (dy+1)*3 + (dx+1))
then b[cout], the order the C ABI uses.

``bf16=True`` rounds what the kernels hold in BF16 (inputs, weights,
stored activations and back-propagated gradients) to BF16 with
round-to-nearest-even, so GPU results can be compared closely.
"""
from __future__ import annotations

import numpy as np

IN_C, OUT_C = 8, 5
OFFSETS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def bf16_round(a):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def layer_shapes(hidden):
    return [(IN_C, hidden), (hidden, hidden), (hidden, OUT_C)]


def param_count(hidden):
    return sum(27 * ci * co + co for ci, co in layer_shapes(hidden))


def unpack(params, hidden):
    """Flat parameter vector -> [(W[27][cin][cout], b[cout])] per layer."""
    out, o = [], 0
    for ci, co in layer_shapes(hidden):
        w = params[o:o + 27 * ci * co].reshape(27, ci, co)
        o += 27 * ci * co
        b = params[o:o + co]
        o += co
        out.append((w, b))
    assert o == len(params)
    return out


def shifted(a, d):
    """a[z + dz, y + dy, x + dx] with periodic wrap."""
    dz, dy, dx = d
    return np.roll(a, shift=(-dz, -dy, -dx), axis=(0, 1, 2))


def conv(a, w, b=None):
    """Periodic 3x3x3 convolution: out[c] = sum_off a[c + d(off)] @ w[off] (+ b)."""
    out = np.zeros(a.shape[:3] + (w.shape[2],))
    for k, d in enumerate(OFFSETS):
        out += shifted(a, d) @ w[k]
    if b is not None:
        out += b
    return out


def conv_bruteforce(a, w, b):
    """Per-cell loop form of conv (tiny grids only)."""
    nz, ny, nx, _ = a.shape
    out = np.zeros((nz, ny, nx, w.shape[2]))
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                acc = np.array(b, dtype=np.float64)
                for k, (dz, dy, dx) in enumerate(OFFSETS):
                    acc = acc + a[(z + dz) % nz, (y + dy) % ny, (x + dx) % nx] @ w[k]
                out[z, y, x] = acc
    return out


def forward(params, x, hidden, bf16=False):
    """Network output (nz, ny, nx, 5) and the stored activations."""
    r = bf16_round if bf16 else (lambda v: v)
    (w1, b1), (w2, b2), (w3, b3) = unpack(params, hidden)
    x = r(x)
    h1 = r(np.maximum(conv(x, r(w1), b1), 0.0))
    h2 = r(np.maximum(conv(h1, r(w2), b2), 0.0))
    y = conv(h2, r(w3), b3)
    return y, (x, h1, h2)


def loss(params, x, target, hidden, bf16=False):
    y, _ = forward(params, x, hidden, bf16)
    return float(np.mean((y - target) ** 2))


def loss_and_grad(params, x, target, hidden, bf16=False):
    """Mean-squared loss and its gradient with respect to the flat parameters."""
    r = bf16_round if bf16 else (lambda v: v)
    (w1, b1), (w2, b2), (w3, b3) = unpack(params, hidden)
    y, (x, h1, h2) = forward(params, x, hidden, bf16)
    V = x.shape[0] * x.shape[1] * x.shape[2]
    L = float(np.mean((y - target) ** 2))
    dy = r(2.0 * (y - target) / (OUT_C * V))

    def wgrad(a, g):
        return np.stack([np.tensordot(shifted(a, d), g, axes=([0, 1, 2], [0, 1, 2])) for d in OFFSETS])

    def input_grad(g, w):
        out = np.zeros(g.shape[:3] + (w.shape[1],))
        for k, d in enumerate(OFFSETS):
            out += shifted(g, tuple(-v for v in d)) @ w[k].T
        return out

    gw3, gb3 = wgrad(h2, dy), dy.sum(axis=(0, 1, 2))
    dh2 = r(input_grad(dy, r(w3))) * (h2 > 0)
    gw2, gb2 = wgrad(h1, dh2), dh2.sum(axis=(0, 1, 2))
    dh1 = r(input_grad(dh2, r(w2))) * (h1 > 0)
    gw1, gb1 = wgrad(x, dh1), dh1.sum(axis=(0, 1, 2))
    grad = np.concatenate([gw1.ravel(), gb1, gw2.ravel(), gb2, gw3.ravel(), gb3])
    return L, grad


def he_init(hidden, seed):
    """Reference initialisation for CPU tests (any seeded scheme will do)."""
    rng = np.random.default_rng(seed)
    parts = []
    for li, (ci, co) in enumerate(layer_shapes(hidden)):
        sd = np.sqrt(2.0 / (27 * ci)) * (0.1 if li == 2 else 1.0)
        parts += [rng.normal(0.0, sd, 27 * ci * co), np.zeros(co)]
    return np.concatenate(parts)


def box_filter2(a):
    """2x2x2 block average of a grid-shaped (nz, ny, nx, ...) array."""
    nz, ny, nx = a.shape[:3]
    s = a.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2, *a.shape[3:])
    return s.mean(axis=(1, 3, 5))


def adam_reference(w, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, Algorithm 1); returns new (w, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    w = w - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return w, m, v
