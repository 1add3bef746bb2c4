"""Pins of the learned-curvature restatement (oracle/mlp.py), on the CPU.

* GELU is the tanh form: torch's gelu(approximate="tanh") to round-off,
  within 4.8e-4 of the erf form x Phi(x), gelu(-x) = gelu(x) - x, and its
  derivative equals central finite differences.
* The float64 backward of the weighted loss equals central finite
  differences.
* BF16 rounding is round-to-nearest-even on known values.
* Labels are +-1/R (minus for the complemented samples); the targets are
  label - kp and the weights 1 / label^2; the features are the BF16
  differences to the centre, the centre - 1/2 and kp.
* The stencils are the tanh profile of a circle: inverting the centre value
  gives the signed distance to the interface, within the offsets the
  generator draws; the centre values lie mostly in [0.02, 0.98].
* The 5 x 5 plain curvature equals the C oracle's curvature (the reference
  algorithm) on a full grid, and on these stencils it is the curvature of
  the level set through the centre, 1 / (R + offset), to a few per cent.
"""
import numpy as np
import pytest

from oracle import mlp


def test_gelu_is_the_tanh_form():
    import torch
    from scipy.special import erf
    x = np.linspace(-12.0, 12.0, 240001)
    g = mlp._gelu(x)
    t = torch.nn.functional.gelu(torch.from_numpy(x), approximate="tanh").numpy()
    assert np.abs(g - t).max() <= 1e-14
    exact = 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))
    assert np.abs(g - exact).max() < 4.8e-4
    assert np.abs(g - exact).max() > 4.5e-4  # the tanh form, not the erf form
    assert np.allclose(mlp._gelu(-x), g - x, rtol=0, atol=1e-14)
    h = 1e-5
    fd = (mlp._gelu(x + h) - mlp._gelu(x - h)) / (2 * h)
    assert np.allclose(mlp._dgelu(x), fd, rtol=1e-8, atol=1e-9)


def test_float64_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = (0.3 * rng.standard_normal(mlp.NP)).astype(np.float64)
    p[mlp.OW1:mlp.OB1].reshape(mlp.NH, mlp.NIP)[:, mlp.NF:] = 0.0
    x = np.zeros((9, mlp.NIP))
    x[:, :mlp.NF] = rng.uniform(-0.5, 0.5, (9, mlp.NF))
    t = rng.uniform(0.0, 0.5, 9)
    w = rng.uniform(0.5, 40.0, 9)
    _, g = mlp.loss_grad64(p, x, t, w)
    idx = np.concatenate([rng.choice(np.arange(o, e), 8, replace=False) for o, e in
                          ((mlp.OW1, mlp.OW1 + mlp.NF), (mlp.OB1, mlp.OW2), (mlp.OW2, mlp.OB2),
                           (mlp.OB2, mlp.OW3), (mlp.OW3, mlp.OB3))] + [np.array([mlp.OB3])])
    eps = 1e-5  # truncation ~1e-10 relative; at 1e-6 round-off of the loss (~28) reaches 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += eps
        pm[k] -= eps
        fd = (mlp.loss_grad64(pp, x, t, w)[0] - mlp.loss_grad64(pm, x, t, w)[0]) / (2 * eps)
        assert fd == pytest.approx(g[k], rel=1e-6, abs=1e-9), k
    # unit weights are the plain mean squared error
    y = mlp.predict64(p, x)
    assert mlp.loss_grad64(p, x, t)[0] == pytest.approx(np.mean((y - t) ** 2), rel=1e-14)


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -7, -2.5, 1.0 + 2.0 ** -9], np.float32)
    want = np.array([1.0, 1.0, 1.0 + 2.0 ** -6, 1.0 + 2.0 ** -7, -2.5, 1.0], np.float32)
    assert np.array_equal(mlp.bf16(x), want)


def test_bf16_path_close_to_float64():
    p = mlp.glorot(11)
    x, _, _, t, w = mlp.samples(5, 0, 256, with_targets=True)
    l64, g64 = mlp.loss_grad64(p, x, t, w)
    l16, g16 = mlp.loss_grad_bf16(p, x, t, w)
    assert l16 == pytest.approx(l64, rel=2e-2)
    assert np.linalg.norm(g16 - g64) <= 5e-2 * np.linalg.norm(g64)


def test_labels_targets_and_features():
    n = 4000
    x, lab, st, t, w = mlp.samples(7, 0, n, with_targets=True)
    i = np.arange(n, dtype=np.uint64)
    R = 6.0 + 58.0 * mlp.uniform(7, i, 1)
    sign = np.where(mlp.uniform(7, i, 11) < 0.5, -1.0, 1.0)
    assert np.allclose(lab, (sign / R).astype(np.float32), rtol=1e-6)
    c = mlp.canonical(st)
    kp = mlp.plain_curvature(c)
    assert np.allclose(kp, mlp.plain_curvature(st), rtol=1e-12, atol=1e-15)
    assert np.allclose(t, (sign / R - kp).astype(np.float32), rtol=1e-6, atol=1e-9)
    assert np.allclose(w, (R * R).astype(np.float32), rtol=1e-6)
    f = np.zeros((n, mlp.NIP))
    f[:, :25] = c - c[:, 12:13]
    f[:, 12] = c[:, 12] - 0.5
    f[:, 25] = kp
    assert np.array_equal(x, mlp.bf16(f.astype(np.float32)))


def d4_images(st):
    """The 8 images of [n][25] stencils under the grid's mirrors and
    transpose."""
    v = np.asarray(st).reshape(-1, 5, 5)
    out = []
    for t in (False, True):
        a = np.swapaxes(v, 1, 2) if t else v
        for mx in (False, True):
            for my in (False, True):
                b = a[:, :, ::-1] if mx else a
                b = b[:, ::-1, :] if my else b
                out.append(b.reshape(-1, 25))
    return out


def test_canonical_orientation_is_the_same_for_every_image():
    """Every mirror / transpose image of a stencil has the same canonical
    stencil, bit for bit (unless its core gradient lies on a symmetry line,
    where the images are themselves symmetric), and it has gx >= gy >= 0;
    the plain curvature of all 8 images agrees to rounding."""
    _, _, st = mlp.samples(23, 0, 2000)
    rng = np.random.default_rng(2)
    st = np.concatenate([st, rng.uniform(0, 1, (500, 25))])
    base = mlp.canonical(st)
    g = base.reshape(-1, 5, 5)
    gx = (g[:, 1:4, 3] - g[:, 1:4, 1]).sum(1)
    gy = (g[:, 3, 1:4] - g[:, 1, 1:4]).sum(1)
    assert np.all(gx >= 0) and np.all(gy >= 0) and np.all(gx >= gy)
    kp = mlp.plain_curvature(st)
    for img in d4_images(st):
        assert np.array_equal(mlp.canonical(img), base)
        assert np.allclose(mlp.plain_curvature(img), kp, rtol=1e-12, atol=1e-14)


def test_stencils_are_the_circle_profile():
    """1/2 (1 + tanh(s (R - d))) inverted at the centre cell gives R - d,
    the signed distance of the centre cell's centre to the interface; it is
    the generator's normal offset (up to 2/s cells) give or take the half
    diagonal of a cell.  The complement mirrors the profile."""
    n = 4000
    seed = 13
    _, lab, st = mlp.samples(seed, 0, n)
    i = np.arange(n, dtype=np.uint64)
    s = 0.5 * 2.0 ** mlp.uniform(seed, i, 8)
    off = (2.0 * mlp.uniform(seed, i, 10) - 1.0) * 2.0 / s
    c = np.where(lab > 0, st[:, 12], 1.0 - st[:, 12])
    dist = np.arctanh(np.clip(2 * c - 1, -1 + 1e-15, 1 - 1e-15)) / s  # R - d
    assert np.all(np.abs(dist + off) <= np.sqrt(0.5) + 1e-9)
    cc = np.minimum(c, 1.0 - c)
    assert cc.min() > 0.002
    assert np.mean(cc >= 0.02) > 0.9


def test_plain_curvature_is_the_level_set_curvature():
    """On the slope-1/2 profile of the configured initial conditions the
    plain curvature of the centre cell is within a few per cent of the
    curvature of the circle through the centre cell's centre, 1/d: it is
    the interface's 1/R only at the interface itself, which is what the
    network corrects."""
    n = 4000
    seed = 17
    _, lab, st = mlp.samples(seed, 0, n)
    i = np.arange(n, dtype=np.uint64)
    s = 0.5 * 2.0 ** mlp.uniform(seed, i, 8)
    R = 6.0 + 58.0 * mlp.uniform(seed, i, 1)
    c = np.where(lab > 0, st[:, 12], 1.0 - st[:, 12])
    d = R - np.arctanh(np.clip(2 * c - 1, -1 + 1e-15, 1 - 1e-15)) / s
    kp = np.sign(lab) * mlp.plain_curvature(st)
    gentle = (s < 0.6) & (R > 12)
    assert np.median(np.abs(kp[gentle] * d[gentle] - 1)) < 0.03
    assert np.mean(np.abs(kp[gentle] * R[gentle] - 1)) > 2 * np.median(np.abs(kp[gentle] * d[gentle] - 1))


def test_plain_curvature_matches_c_oracle():
    import oracle
    from simplebubblyflowsolver_amd import ic
    nx = ny = 48
    o = oracle.new_case(nx, ny, 1.0, 1.0)
    try:
        vof = ic.bubble_vof(nx, ny, 1.0, 1.0, 0.5, 0.5, 0.2)
        o.load_initial_condition(vof)
        o.compute_curvature_tensor()
        curv = o.get("curv")           # rows j = 0..ny+1, cols i = 0..nx+1
        v = o.get("vof")               # rows j = -1..ny+2
        h = 1.0 / nx
        sts, want = [], []
        for j in range(3, ny - 2):
            for i in range(3, nx - 2):
                sts.append(v[j + 1 - 2:j + 1 + 3, i - 2:i + 3].ravel())
                want.append(curv[j, i] * h)
        got = mlp.plain_curvature(np.array(sts))
        assert np.allclose(got, np.array(want), rtol=1e-12, atol=1e-14)
    finally:
        o.close()
