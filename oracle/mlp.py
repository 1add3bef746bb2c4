"""TEST INFRASTRUCTURE ONLY: CPU restatement of the learned-curvature pieces.

Only tests/ may import this module; the product never does.

The network has no counterpart in the reference. What it replaces is the
plain curvature of src/interface/curvature_tensor.c:16-135, whose sign
convention (kappa > 0 with the gas, vof = 1, inside) the labels follow.

* `uniform`: the counter-based generator (splitmix64 finaliser) the
  device uses.
* `samples`: the training stencils, restated from their definition: the
  initial-condition profile 1/2 (1 + tanh(s (R - d))) around circles
  (initial_condition/main.py:40-41) at random sub-cell offsets, slopes and
  normal offsets, in the 8 symmetries; half are complemented (1 - vof, label
  -kappa h). The tests pin this by closed forms: the label is +-1/R, and the
  values are the formula at the cell centres.
* The network corrects the plain curvature: its features are the stencil
  in canonical orientation (`canonical`: of its 8 images under the grid's
  mirrors and transpose, the one whose core gradient has gx >= gy >= 0) and
  kp = `plain_curvature` of it; its output y is added to kp, and training
  minimises mean(w (y - t)^2) with t = label - kp, w = 1 / label^2.
* `plain_curvature`: the reference's curvature of the centre cell of a
  stencil with h = 1 (curvature_tensor.c:16-49 corner gradients, :100-135
  divergence). The tests pin it against the C oracle on a full grid.
* `_gelu`: GELU in Hendrycks & Gimpel's tanh form, pinned against torch's
  gelu(approximate="tanh"), the erf form x Phi(x) (within 4.8e-4) and the
  identity gelu(-x) = gelu(x) - x.
* `loss_grad64`: forward and backward of the weighted loss in float64 with
  no rounding, pinned by central finite differences.
* `loss_grad_bf16`: the same with the device kernels' BF16 roundings (inputs,
  weights, layer-1 activations and the deltas that feed MFMA operands). This
  is the checker of the device gradients.
"""
from __future__ import annotations

import numpy as np

NI, NF, NIP, NH = 25, 26, 32, 64  # stencil values, features used, padded, hidden
OW1 = 0
OB1 = OW1 + NH * NIP
OW2 = OB1 + NH
OB2 = OW2 + NH * NH
OW3 = OB2 + NH
OB3 = OW3 + NH
NP = OB3 + 1

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _mix64(z: np.ndarray) -> np.ndarray:
    z = z + np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def uniform(seed: int, i: np.ndarray, k: int) -> np.ndarray:
    """Uniform doubles in [0, 1) of samples i, stream k."""
    with np.errstate(over="ignore"):
        key = (np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15)
               + np.asarray(i, dtype=np.uint64) * np.uint64(16) + np.uint64(k))
        return (_mix64(key) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def bf16(x) -> np.ndarray:
    """Round float32 values to BF16 (nearest even), returned as float32."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + np.uint64(0x7FFF) + ((u >> np.uint64(16)) & np.uint64(1))) >> np.uint64(16)
    return (u.astype(np.uint32) << np.uint32(16)).view(np.float32)


def canonical(st: np.ndarray) -> np.ndarray:
    """Each [25] stencil in the orientation, among its 8 images under the
    grid's mirrors and transpose, whose 3 x 3 core gradient sums (gx, gy)
    satisfy gx >= gy >= 0."""
    v = np.asarray(st, dtype=np.float64).reshape(-1, 25)
    gx = np.zeros(v.shape[0])
    gy = np.zeros(v.shape[0])
    for r in (1, 2, 3):
        gx = gx + (v[:, r * 5 + 3] - v[:, r * 5 + 1])
    for c in (1, 2, 3):
        gy = gy + (v[:, 3 * 5 + c] - v[:, 1 * 5 + c])
    sx = np.where(gx < 0, -1, 1)
    sy = np.where(gy < 0, -1, 1)
    tr = np.abs(gy) > np.abs(gx)
    out = np.empty_like(v)
    for dj in range(-2, 3):
        for di in range(-2, 3):
            si = np.where(tr, sx * dj, sx * di)
            sj = np.where(tr, sy * di, sy * dj)
            out[:, (dj + 2) * 5 + di + 2] = v[np.arange(v.shape[0]), (sj + 2) * 5 + si + 2]
    return out


def features(st: np.ndarray) -> np.ndarray:
    """[n][25] stencils -> [n][32] float32 features before BF16 rounding, of
    the canonical stencil: the differences to the centre, the centre - 1/2,
    and kp."""
    st = canonical(st)
    f = st - st[:, 12:13]
    f[:, 12] = st[:, 12] - 0.5
    out = np.zeros((st.shape[0], NIP), np.float32)
    out[:, :NI] = f.astype(np.float32)
    out[:, NI] = plain_curvature(st).astype(np.float32)
    return out


def targets(labels64: np.ndarray, st: np.ndarray):
    """Training targets label - kp (kp of the canonical stencil) and weights
    1 / label^2 (float32) from float64 labels."""
    lab = np.asarray(labels64, dtype=np.float64)
    return (lab - plain_curvature(canonical(st))).astype(np.float32), (1.0 / (lab * lab)).astype(np.float32)


def samples(seed: int, first: int, n: int, with_targets: bool = False):
    """(features [n][32] BF16 values, labels [n], stencils [n][25]), and the
    targets and weights if `with_targets`."""
    i = np.arange(first, first + n, dtype=np.uint64)
    u = {k: uniform(seed, i, k) for k in (1, 4, 5, 6, 8, 9, 10, 11)}
    b = 6.0 + 58.0 * u[1]
    t = 2.0 * np.pi * u[4]
    ox, oy = u[5], u[6]
    slope = 0.5 * 2.0 ** u[8]
    sym = (u[9] * 8.0).astype(np.int64)
    off = (2.0 * u[10] - 1.0) * 2.0 / slope
    px, py = (b + off) * np.cos(t), (b + off) * np.sin(t)
    comp = u[11] < 0.5
    label = np.where(comp, -1.0, 1.0) / b
    ic, jc = np.floor(px - ox), np.floor(py - oy)
    v = np.zeros((n, NI))
    for dj in range(-2, 3):
        for di in range(-2, 3):
            cx, cy = ic + di + ox + 0.5, jc + dj + oy + 0.5
            val = 0.5 * (1.0 + np.tanh(slope * (b - np.sqrt(cx * cx + cy * cy))))
            ti, tj = np.full(n, di), np.full(n, dj)
            t1 = (sym & 1) != 0
            ti, tj = np.where(t1, tj, ti), np.where(t1, ti, tj)
            ti = np.where((sym & 2) != 0, -ti, ti)
            tj = np.where((sym & 4) != 0, -tj, tj)
            v[np.arange(n), (tj + 2) * 5 + (ti + 2)] = np.where(comp, 1.0 - val, val)
    out = (bf16(features(v)), label.astype(np.float32), v)
    if with_targets:
        out = out + targets(label, v)
    return out


def plain_curvature(st: np.ndarray) -> np.ndarray:
    """kappa h of the centre cell of [n][25] stencils, as the reference
    computes it on a uniform grid of cells h x h."""
    v = np.asarray(st, dtype=np.float64).reshape(-1, 5, 5)

    def corner(ai, bj):
        mm, pm = v[:, bj - 1, ai - 1], v[:, bj - 1, ai]
        mp, pp = v[:, bj, ai - 1], v[:, bj, ai]
        gx = 1.0 / 1.0 * (-mm + pm - mp + pp)
        gy = 1.0 / 1.0 * (-mm - pm + mp + pp)
        ninv = 1.0 / np.maximum(np.sqrt(gx * gx + gy * gy), np.finfo(np.float64).eps)
        return gx * ninv, gy * ninv

    bl, br, tl, tr = corner(2, 2), corner(3, 2), corner(2, 3), corner(3, 3)
    nxm = 0.5 * bl[0] + 0.5 * tl[0]
    nxp = 0.5 * br[0] + 0.5 * tr[0]
    nym = 0.5 * bl[1] + 0.5 * br[1]
    nyp = 0.5 * tl[1] + 0.5 * tr[1]
    return -1.0 * (-nxm + nxp - nym + nyp)


def glorot(seed: int) -> np.ndarray:
    """The device's initial parameters."""
    p = np.zeros(NP, np.float32)
    r1, r2, r3 = np.sqrt(6.0 / (NF + NH)), np.sqrt(6.0 / (NH + NH)), np.sqrt(6.0 / (NH + 1))
    o, k = np.meshgrid(np.arange(NH), np.arange(NF), indexing="ij")
    idx = (1000000000 + o * NIP + k).astype(np.uint64)
    p[OW1 + o * NIP + k] = (r1 * (2.0 * uniform(seed, idx, 0) - 1.0)).astype(np.float32)
    p[OW2:OW2 + NH * NH] = (r2 * (2.0 * uniform(seed, 2000000000 + np.arange(NH * NH), 0) - 1.0)).astype(np.float32)
    p[OW3:OW3 + NH] = (r3 * (2.0 * uniform(seed, 3000000000 + np.arange(NH), 0) - 1.0)).astype(np.float32)
    return p


# GELU in the tanh form of Hendrycks & Gimpel (torch's gelu(approximate="tanh")):
# x (1 + tanh u) / 2, u = sqrt(2/pi) (x + 0.044715 x^3)
_GK0 = np.sqrt(2.0 / np.pi)
_GK1 = 0.044715 * _GK0


def _gelu(x):
    return 0.5 * x * (1.0 + np.tanh(x * (_GK0 + _GK1 * x * x)))


def _dgelu(x):
    th = np.tanh(x * (_GK0 + _GK1 * x * x))
    return 0.5 * (1.0 + th) + 0.5 * x * (1.0 - th * th) * (_GK0 + 3.0 * _GK1 * x * x)


def _unpack(p):
    p = np.asarray(p)
    return (p[OW1:OB1].reshape(NH, NIP), p[OB1:OW2], p[OW2:OB2].reshape(NH, NH), p[OB2:OW3],
            p[OW3:OB3], p[OB3])


def _pack(gw1, gb1, gw2, gb2, gw3, gb3):
    g = np.zeros(NP, np.float64)
    g[OW1:OB1] = gw1.ravel()
    g[OB1:OW2] = gb1
    g[OW2:OB2] = gw2.ravel()
    g[OB2:OW3] = gb2
    g[OW3:OB3] = gw3
    g[OB3] = gb3
    return g


def predict64(p, x):
    w1, b1, w2, b2, w3, b3 = (np.asarray(a, np.float64) for a in _unpack(p))
    x = np.asarray(x, np.float64)
    h1 = w1 @ x.T + b1[:, None]
    h2 = w2 @ _gelu(h1) + b2[:, None]
    return w3 @ _gelu(h2) + b3


def loss_grad64(p, x, t, wt=None):
    """Weighted mean squared error mean(wt (y - t)^2) and its gradient,
    float64, no rounding (wt None: all 1)."""
    w1, b1, w2, b2, w3, b3 = (np.asarray(a, np.float64) for a in _unpack(p))
    x, t = np.asarray(x, np.float64), np.asarray(t, np.float64)
    n = x.shape[0]
    wt = np.ones(n) if wt is None else np.asarray(wt, np.float64)
    h1 = w1 @ x.T + b1[:, None]
    a1 = _gelu(h1)
    h2 = w2 @ a1 + b2[:, None]
    a2 = _gelu(h2)
    y = w3 @ a2 + b3
    e = y - t
    dy = 2.0 / n * wt * e
    d2 = w3[:, None] * dy[None, :] * _dgelu(h2)
    d1 = (w2.T @ d2) * _dgelu(h1)
    return float(np.mean(wt * e * e)), _pack(d1 @ x, d1.sum(1), d2 @ a1.T, d2.sum(1), a2 @ dy, dy.sum())


def loss_grad_bf16(p, x, t, wt=None):
    """The same with the device's BF16 operand roundings (x must already be
    BF16 values); sums in float64, activations in float32."""
    w1, b1, w2, b2, w3, b3 = _unpack(np.asarray(p, np.float32))
    w1b, w2b = bf16(w1).astype(np.float64), bf16(w2).astype(np.float64)
    x = np.asarray(x, np.float32).astype(np.float64)
    t = np.asarray(t, np.float64)
    n = x.shape[0]
    wt = np.ones(n, np.float32) if wt is None else np.asarray(wt, np.float32)
    f32 = lambda a: np.asarray(a, np.float64).astype(np.float32).astype(np.float64)  # noqa: E731
    h1 = f32(w1b @ x.T + b1[:, None])
    a1b = bf16(f32(_gelu(h1))).astype(np.float64)
    h2 = f32(w2b @ a1b + b2[:, None])
    a2 = f32(_gelu(h2))
    y = f32(w3.astype(np.float64) @ a2 + b3)
    e = f32(y - t)
    we = f32(wt.astype(np.float64) * e)
    dy = f32(2.0 / n * we)
    d2 = f32(w3[:, None] * dy[None, :] * f32(_dgelu(h2)))
    d2b = bf16(d2).astype(np.float64)
    d1 = f32(f32(w2b.T @ d2b) * f32(_dgelu(h1)))
    d1b = bf16(d1).astype(np.float64)
    return float(np.mean(we * e)), _pack(d1b @ x, d1.sum(1), d2b @ a1b.T, d2.sum(1), a2 @ dy, dy.sum())
