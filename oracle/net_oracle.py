"""CPU oracle of the learned warm start's network -- TEST INFRASTRUCTURE.

Only tests/ (and __graft_entry__.smoke) use this module, as the checker of the
library's HIP kernels; nothing in the product path imports it.  The network has
no counterpart in the reference: it is pinned here by central finite
differences of its own loss (tests/test_net_oracle.py).

Network (per pixel of an H x W image, channels last, one-pixel zero border):
  X0 (8) -conv3x3-> X1 (32, ReLU) -conv3x3-> X2 (32, ReLU) -conv3x3-> X3 (32, ReLU)
     -conv3x3-> delta (1)
  loss = mean over pixels of (delta - target)^2
delta is 0 on pixels whose 3x3 neighbourhood has |feature 5| <= 1e-5 (no
activity), and those pixels carry no gradient.
Parameter vector (float32), in order: W1[32][9][8], b1[32], W2[32][9][32],
b2[32], W3[32][9][32], b3[32], W4[9][32], b4[1]; tap = 3 (dy + 1) + (dx + 1).

With bf16=True the stored quantities are rounded to bfloat16 where the HIP
kernels store them in bfloat16: the weights of the four layers, the
activations X1..X3 and the backward-pass gradients of the pre-activations of
layers 1..3.  Accumulation is float64 here (float32 on the GPU).
"""
from __future__ import annotations

import numpy as np

C0, C = 8, 32
SIZES = [("W1", C * 9 * C0), ("b1", C), ("W2", C * 9 * C), ("b2", C), ("W3", C * 9 * C), ("b3", C),
         ("W4", 9 * C), ("b4", 1)]
NUM_PARAMS = sum(n for _, n in SIZES)
TAPS = [(t // 3 - 1, t % 3 - 1) for t in range(9)]


def bf16_round(a):
    """Round float32 values to the nearest bfloat16 (ties to even), as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    out = r.view(np.float32).copy()
    nan = np.isnan(a)
    out[nan] = np.nan
    return out


def bf16_bits(a):
    """bfloat16 bit patterns (uint16) of float32 values (rounded to nearest even)."""
    return (bf16_round(a).view(np.uint32) >> 16).astype(np.uint16)


def split(p):
    out, off = {}, 0
    for name, n in SIZES:
        out[name] = np.asarray(p[off : off + n], dtype=np.float64)
        off += n
    out["W1"] = out["W1"].reshape(C, 9, C0)
    out["W2"] = out["W2"].reshape(C, 9, C)
    out["W3"] = out["W3"].reshape(C, 9, C)
    out["W4"] = out["W4"].reshape(9, C)
    return out


def join(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SIZES])


def _rnd(a, bf16):
    return bf16_round(a.astype(np.float32)).astype(np.float64) if bf16 else a


def _conv(Xp, W, H, Wd):
    """Xp: (H+2, W+2, cin) padded; W: (cout, 9, cin) -> (H, W, cout)."""
    out = np.zeros((H, Wd, W.shape[0]))
    for t, (dy, dx) in enumerate(TAPS):
        out += Xp[1 + dy : 1 + dy + H, 1 + dx : 1 + dx + Wd, :] @ W[:, t, :].T
    return out


def _pad(X):
    H, Wd, c = X.shape
    P = np.zeros((H + 2, Wd + 2, c))
    P[1:-1, 1:-1] = X
    return P


def forward(params, X0p, bf16=True, gate=None):
    """X0p: (H+2, W+2, 8) padded features.  Returns delta (H, W) and the padded
    activations [X0p, X1p, X2p, X3p]."""
    P = split(params)
    H, Wd = X0p.shape[0] - 2, X0p.shape[1] - 2
    acts = [np.asarray(X0p, dtype=np.float64)]
    for l in (1, 2, 3):
        W = _rnd(P[f"W{l}"], bf16)
        pre = _conv(acts[-1], W, H, Wd) + P[f"b{l}"]
        acts.append(_pad(_rnd(np.maximum(pre, 0.0), bf16)))
    W4 = _rnd(P["W4"], bf16)
    delta = _conv(acts[-1], W4[None], H, Wd)[..., 0] + P["b4"][0]
    return delta * active(X0p, gate), acts


GATE = 1e-5  # the library default (cdiff_net_set_gate)


def active(X0p, gate=None):
    """Pixels whose 3x3 neighbourhood has |feature 5| > gate (the prediction is
    0 elsewhere)."""
    gate = GATE if gate is None else gate
    H, Wd = X0p.shape[0] - 2, X0p.shape[1] - 2
    e = np.zeros((H, Wd))
    for dy, dx in TAPS:
        e = np.maximum(e, np.abs(X0p[1 + dy : 1 + dy + H, 1 + dx : 1 + dx + Wd, 5]))
    return (e > gate).astype(np.float64)


def loss_grad(params, X0p, target, bf16=True, gate=None):
    """(loss, gradient vector, delta)."""
    P = split(params)
    H, Wd = X0p.shape[0] - 2, X0p.shape[1] - 2
    delta, acts = forward(params, X0p, bf16, gate)
    npx = H * Wd
    d = delta - target
    loss = float(np.sum(d * d) / npx)
    G4 = 2.0 * d / npx * active(X0p, gate)
    grads = {}
    # head
    W4 = _rnd(P["W4"], bf16)
    X3p = acts[3]
    grads["W4"] = np.stack([np.einsum("yx,yxc->c", G4, X3p[1 + dy : 1 + dy + H, 1 + dx : 1 + dx + Wd])
                            for dy, dx in TAPS])
    grads["b4"] = np.array([G4.sum()])
    dX = np.zeros_like(X3p)
    for t, (dy, dx) in enumerate(TAPS):
        dX[1 + dy : 1 + dy + H, 1 + dx : 1 + dx + Wd] += G4[..., None] * W4[t][None, None, :]
    G = _rnd(dX[1:-1, 1:-1] * (X3p[1:-1, 1:-1] > 0), bf16)
    for l in (3, 2, 1):
        Xin = acts[l - 1]
        W = _rnd(P[f"W{l}"], bf16)
        grads[f"W{l}"] = np.stack(
            [np.einsum("yxo,yxi->oi", G, Xin[1 + dy : 1 + dy + H, 1 + dx : 1 + dx + Wd]) for dy, dx in TAPS], axis=1)
        grads[f"b{l}"] = G.sum(axis=(0, 1))
        if l > 1:
            dX = np.zeros_like(Xin)
            for t, (dy, dx) in enumerate(TAPS):
                dX[1 + dy : 1 + dy + H, 1 + dx : 1 + dx + Wd] += G @ W[:, t, :]
            G = _rnd(dX[1:-1, 1:-1] * (Xin[1:-1, 1:-1] > 0), bf16)
    return loss, join(grads), delta


def init_params(seed):
    rng = np.random.default_rng(seed)
    d = {
        "W1": rng.normal(0, np.sqrt(2 / 54), (C, 9, C0)),
        "b1": rng.normal(0, 0.1, C),
        "W2": rng.normal(0, np.sqrt(2 / (9 * C)), (C, 9, C)),
        "b2": rng.normal(0, 0.1, C),
        "W3": rng.normal(0, np.sqrt(2 / (9 * C)), (C, 9, C)),
        "b3": rng.normal(0, 0.1, C),
        "W4": rng.normal(0, np.sqrt(1 / (9 * C)), (9, C)),
        "b4": rng.normal(0, 0.1, 1),
    }
    return join(d).astype(np.float32)


def random_features(H, W, seed):
    """Padded (H+2, W+2, 8) float32 features: 6 random channels, a ones mask, a zero channel."""
    rng = np.random.default_rng(seed)
    X = np.zeros((H + 2, W + 2, C0), dtype=np.float32)
    X[1:-1, 1:-1, :6] = rng.normal(0, 1, (H, W, 6))
    X[1:-1, 1:-1, 6] = 1.0
    return X
