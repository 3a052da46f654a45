"""CPU oracle of the learned warm start's network -- TEST INFRASTRUCTURE.

Only tests/ (and __graft_entry__.smoke) use this module, as the checker of the
library's HIP kernels; nothing in the product path imports it.  The network has
no counterpart in the reference: it is pinned here by central finite
differences of its own loss (tests/test_net_oracle.py).

Network: a 4-level encoder-decoder.  Level 0 is the H x W image of cells;
level l = 1..3 has ceil(H_{l-1} / 2) rows and ceil(W_{l-1} / 2) columns.
  XS = S2D(X0)                  level 1: channel 8 (2 dy + dx) + f of pixel
                                (y, x) is feature f of cell (2y + dy, 2x + dx);
                                feature 7 is the largest |feature 5| over the
                                cell's 3x3 neighbourhood (its activity)
  E1 = relu(conv1(XS))          level 1
  E2 = relu(conv2(pool(E1)))    level 2    pool: mean of each 2x2 block
  E3 = relu(conv3(pool(E2)))    level 3
  B3 = relu(conv4(E3))          level 3
  U2 = relu(conv5(up(B3) + E2)) level 2    up: each pixel to its 2x2 block
  U1 = relu(conv6(up(U2) + E1)) level 1
  delta(2y + dy, 2x + dx) = head(U1)(y, x)[2 dy + dx]
conv1..conv6: 3x3, 32 -> 32 with bias; head: 3x3, 32 -> 4 with bias; pixels
beyond an image edge (the one-pixel border, the cells beyond an odd edge)
are 0.  delta is 0 on cells whose 3x3 neighbourhood has |feature 5| <= gate
(no activity), and those cells carry no gradient.

Loss (kind 0): the mean squared error of the normalised increment.  Loss
(kind 1): the linearised residual of the prediction's error relative to
that of the start from T^n.  With e = delta - target and D_f = expm1(feature 3) on a
cell's east face, expm1(feature 4) on its north face (the face diffusion
numbers dt abar g; west and south faces take the neighbour's), the operator
(A v)(c) = v(c) + sum over the faces of c with a neighbour inside the image
of D_f (v(c) - v(nbr)) -- the implicit step I - dt L with L the orthogonal
part of the diffusion operator -- gives loss = |A e|^2 / |A target|^2 (0
when A target = 0): the squared ratio of the FP64 residuals of the predicted
start and of T^n, to first order.
Parameter vector (float32), in order: W1[32][9][32], b1[32], ..., W6, b6,
WH[4][9][32], bH[4]; tap = 3 (dy + 1) + (dx + 1).

With bf16=True the stored quantities are rounded to bfloat16 where the HIP
kernels store them in bfloat16: the weights, the activations E1, P1 =
pool(E1), E2, P2, E3, B3, A5 = up(B3) + E2, U2, A6 = up(U2) + E1, U1, and the
backward pass's gradients of U1, A6, U2, A5, B3, E3, P2, E2, P1, E1.
Accumulation is float64 here (float32 on the GPU).
"""
from __future__ import annotations

import numpy as np

C0, C, HO = 8, 32, 4
SIZES = [item for l in range(1, 7) for item in ((f"W{l}", C * 9 * C), (f"b{l}", C))] + [("WH", HO * 9 * C),
                                                                                      ("bH", HO)]
NUM_PARAMS = sum(n for _, n in SIZES)
TAPS = [(t // 3 - 1, t % 3 - 1) for t in range(9)]
GATE = 1e-5  # the library default (cdiff_net_set_gate)


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
    for l in range(1, 7):
        out[f"W{l}"] = out[f"W{l}"].reshape(C, 9, C)
    out["WH"] = out["WH"].reshape(HO, 9, C)
    return out


def join(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SIZES])


def _rnd(a, bf16):
    return bf16_round(a.astype(np.float32)).astype(np.float64) if bf16 else a


def _half(n):
    return (n + 1) // 2


def _pad(X):
    H, Wd, c = X.shape
    P = np.zeros((H + 2, Wd + 2, c))
    P[1:-1, 1:-1] = X
    return P


def _conv(Xp, W):
    """Xp: (H+2, W+2, cin) padded; W: (cout, 9, cin) -> (H, W, cout)."""
    H, Wd = Xp.shape[0] - 2, Xp.shape[1] - 2
    out = np.zeros((H, Wd, W.shape[0]))
    for t, (dy, dx) in enumerate(TAPS):
        out += Xp[1 + dy : 1 + dy + H, 1 + dx : 1 + dx + Wd, :] @ W[:, t, :].T
    return out


def _conv_t(G, W):
    """Adjoint of _conv in its input: G (H, W, cout) -> (H, W, cin)."""
    H, Wd = G.shape[:2]
    dX = np.zeros((H + 2, Wd + 2, W.shape[2]))
    for t, (dy, dx) in enumerate(TAPS):
        dX[1 + dy : 1 + dy + H, 1 + dx : 1 + dx + Wd] += G @ W[:, t, :]
    return dX[1:-1, 1:-1]


def _wgrad(G, Xp):
    """Gradient of _conv in W: (cout, 9, cin)."""
    H, Wd = G.shape[:2]
    return np.stack([np.einsum("yxo,yxi->oi", G, Xp[1 + dy : 1 + dy + H, 1 + dx : 1 + dx + Wd]) for dy, dx in TAPS],
                    axis=1)


def _even(X):
    """X (H, W, c) zero-extended to even H and W."""
    H, Wd, c = X.shape
    E = np.zeros((H + H % 2, Wd + Wd % 2, c))
    E[:H, :Wd] = X
    return E


def _pool(X):
    """Mean of the 2x2 blocks of X (H, W, c): ((a + b) + (d + f)) / 4."""
    E = _even(X)
    return ((E[0::2, 0::2] + E[0::2, 1::2]) + (E[1::2, 0::2] + E[1::2, 1::2])) * 0.25


def _up(Cc, H, Wd):
    """Each pixel of Cc to its 2x2 block, cut to H x W."""
    return np.repeat(np.repeat(Cc, 2, axis=0), 2, axis=1)[:H, :Wd]


def _up_t(G):
    """Adjoint of _up: the sum over each 2x2 block."""
    E = _even(G)
    return (E[0::2, 0::2] + E[0::2, 1::2]) + (E[1::2, 0::2] + E[1::2, 1::2])


def activity(X0p):
    """Feature 7 of every cell: the largest |feature 5| over its 3x3
    neighbourhood (the zero border included)."""
    H, Wd = X0p.shape[0] - 2, X0p.shape[1] - 2
    e = np.zeros((H, Wd))
    for dy, dx in TAPS:
        e = np.maximum(e, np.abs(X0p[1 + dy : 1 + dy + H, 1 + dx : 1 + dx + Wd, 5]))
    return e


def s2d(X0p):
    """(H+2, W+2, 8) padded features -> (H1, W1, 32) level-1 input, feature 7
    replaced by the activity."""
    X = np.array(X0p[1:-1, 1:-1], dtype=np.float64)
    X[..., 7] = activity(X0p)
    E = _even(X)
    return np.concatenate([E[dy::2, dx::2] for dy in (0, 1) for dx in (0, 1)], axis=2)


def d2s(D4, H, Wd):
    """(H1, W1, 4) head outputs -> (H, W) cells."""
    out = np.zeros((2 * D4.shape[0], 2 * D4.shape[1]))
    for dy in (0, 1):
        for dx in (0, 1):
            out[dy::2, dx::2] = D4[..., 2 * dy + dx]
    return out[:H, :Wd]


def active(X0p, gate=None):
    """Cells whose 3x3 neighbourhood has |feature 5| > gate (the prediction is
    0 elsewhere)."""
    gate = GATE if gate is None else gate
    return (activity(X0p) > gate).astype(np.float64)


def forward(params, X0p, bf16=True, gate=None):
    """X0p: (H+2, W+2, 8) padded features.  Returns delta (H, W) and the
    activations (unpadded, keyed by name)."""
    P = split(params)
    H, Wd = X0p.shape[0] - 2, X0p.shape[1] - 2
    W = {k: _rnd(v, bf16) if k.startswith("W") else v for k, v in P.items()}
    a = {"XS": s2d(X0p)}

    def conv(l, Xin):
        return _rnd(np.maximum(_conv(_pad(Xin), W[f"W{l}"]) + P[f"b{l}"], 0.0), bf16)

    H1, W1 = _half(H), _half(Wd)
    H2, W2 = _half(H1), _half(W1)
    a["E1"] = conv(1, a["XS"])
    a["P1"] = _rnd(_pool(a["E1"]), bf16)
    a["E2"] = conv(2, a["P1"])
    a["P2"] = _rnd(_pool(a["E2"]), bf16)
    a["E3"] = conv(3, a["P2"])
    a["B3"] = conv(4, a["E3"])
    a["A5"] = _rnd(_up(a["B3"], H2, W2) + a["E2"], bf16)
    a["U2"] = conv(5, a["A5"])
    a["A6"] = _rnd(_up(a["U2"], H1, W1) + a["E1"], bf16)
    a["U1"] = conv(6, a["A6"])
    D4 = _conv(_pad(a["U1"]), W["WH"]) + P["bH"]
    return d2s(D4, H, Wd) * active(X0p, gate), a


def face_numbers(X0p):
    """(D_east, D_north) per cell, (H, W) each: expm1 of features 3 and 4, 0 on
    faces without a neighbour inside the image."""
    H, Wd = X0p.shape[0] - 2, X0p.shape[1] - 2
    De = np.expm1(np.asarray(X0p[1:-1, 1:-1, 3], dtype=np.float64))
    Dn = np.expm1(np.asarray(X0p[1:-1, 1:-1, 4], dtype=np.float64))
    De[:, Wd - 1] = 0.0
    Dn[H - 1, :] = 0.0
    return De, Dn


def apply_A(v, De, Dn):
    """(A v)(c) = v(c) + sum_f D_f (v(c) - v(nbr)); A is symmetric."""
    r = v.copy()
    fe = De[:, :-1] * (v[:, :-1] - v[:, 1:])  # east face of (y, x): between x and x+1
    r[:, :-1] += fe
    r[:, 1:] -= fe
    fn = Dn[:-1, :] * (v[:-1, :] - v[1:, :])  # north face of (y, x): between y and y+1
    r[:-1, :] += fn
    r[1:, :] -= fn
    return r


def residual_loss(delta, target, X0p, kind=1):
    """(loss, d loss / d delta before the gate, the normaliser): kind 1, |A e|^2
    / |A target|^2; kind 0, the mean squared error |e|^2 / (H W)."""
    if kind == 0:
        e = delta - target
        return float(np.sum(e * e)) / e.size, 2.0 * e / e.size, float(e.size)
    De, Dn = face_numbers(X0p)
    re = apply_A(delta - target, De, Dn)
    b = float(np.sum(apply_A(target, De, Dn) ** 2))
    if b == 0.0:
        return 0.0, np.zeros_like(delta), 0.0
    return float(np.sum(re * re)) / b, 2.0 * apply_A(re, De, Dn) / b, b


def loss_grad(params, X0p, target, bf16=True, gate=None, kind=0):
    """(loss, gradient vector, delta); kind as residual_loss."""
    P = split(params)
    W = {k: _rnd(v, bf16) if k.startswith("W") else v for k, v in P.items()}
    H, Wd = X0p.shape[0] - 2, X0p.shape[1] - 2
    H1, W1 = _half(H), _half(Wd)
    delta, a = forward(params, X0p, bf16, gate)
    loss, gd, _ = residual_loss(delta, target, X0p, kind)
    Gc = _even((gd * active(X0p, gate))[..., None])[..., 0]
    G4 = np.stack([Gc[dy::2, dx::2] for dy in (0, 1) for dx in (0, 1)], axis=2)[:H1, :W1]
    g = {"WH": _wgrad(G4, _pad(a["U1"])), "bH": G4.sum(axis=(0, 1))}

    def r(x):
        return _rnd(x, bf16)

    def wg(l, G, Xin):
        g[f"W{l}"] = _wgrad(G, _pad(Xin))
        g[f"b{l}"] = G.sum(axis=(0, 1))

    GU1 = r(_conv_t(G4, W["WH"]) * (a["U1"] > 0))
    wg(6, GU1, a["A6"])
    GA6 = r(_conv_t(GU1, W["W6"]))
    GU2 = r(_up_t(GA6) * (a["U2"] > 0))
    wg(5, GU2, a["A5"])
    GA5 = r(_conv_t(GU2, W["W5"]))
    GB3 = r(_up_t(GA5) * (a["B3"] > 0))
    wg(4, GB3, a["E3"])
    GE3 = r(_conv_t(GB3, W["W4"]) * (a["E3"] > 0))
    wg(3, GE3, a["P2"])
    GP2 = r(_conv_t(GE3, W["W3"]))
    GE2 = r((GA5 + 0.25 * _up(GP2, *a["E2"].shape[:2])) * (a["E2"] > 0))
    wg(2, GE2, a["P1"])
    GP1 = r(_conv_t(GE2, W["W2"]))
    GE1 = r((GA6 + 0.25 * _up(GP1, *a["E1"].shape[:2])) * (a["E1"] > 0))
    wg(1, GE1, a["XS"])
    return loss, join(g), delta


def init_params(seed):
    rng = np.random.default_rng(seed)
    d = {}
    for l in range(1, 7):
        d[f"W{l}"] = rng.normal(0, np.sqrt(2 / (9 * (24 if l == 1 else C))), (C, 9, C))
        d[f"b{l}"] = rng.normal(0, 0.1, C)
    d["WH"] = rng.normal(0, np.sqrt(1 / (9 * C)), (HO, 9, C))
    d["bH"] = rng.normal(0, 0.1, HO)
    return join(d).astype(np.float32)


def random_features(H, W, seed):
    """Padded (H+2, W+2, 8) float32 features: 6 random channels, a ones mask, a zero channel."""
    rng = np.random.default_rng(seed)
    X = np.zeros((H + 2, W + 2, C0), dtype=np.float32)
    X[1:-1, 1:-1, :6] = rng.normal(0, 1, (H, W, 6))
    X[1:-1, 1:-1, 6] = 1.0
    return X
