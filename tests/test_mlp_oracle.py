"""Pins of the learned-curvature restatement (oracle/mlp.py), on the CPU.

* The float64 backward equals central finite differences of the loss.
* BF16 rounding is round-to-nearest-even on known values.
* Labels: +-1/R for circles (minus for the complemented samples); the
  ellipse formula equals the curvature of the parametric curve by finite
  differences.
* Sharp stencils hold the area of the shape inside each cell: their sum
  equals the area of the shape (or of its complement) inside the 5 x 5
  window, measured by a 16x finer sampling, whatever the symmetry transform
  and the neighbour shift.
* Diffuse stencils are centred where the vof is mostly in [0.02, 0.98].
* The 5 x 5 plain curvature equals the C oracle's curvature (the reference
  algorithm) on a full grid.
"""
import numpy as np
import pytest

from oracle import mlp


def test_float64_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = (0.3 * rng.standard_normal(mlp.NP)).astype(np.float64)
    p[mlp.OW1:mlp.OB1].reshape(mlp.NH, mlp.NIP)[:, mlp.NI:] = 0.0
    x = np.zeros((9, mlp.NIP))
    x[:, :mlp.NI] = rng.uniform(-0.5, 0.5, (9, mlp.NI))
    t = rng.uniform(0.0, 0.5, 9)
    _, g = mlp.loss_grad64(p, x, t)
    idx = np.concatenate([rng.choice(np.arange(o, e), 8, replace=False) for o, e in
                          ((mlp.OW1, mlp.OW1 + mlp.NI), (mlp.OB1, mlp.OW2), (mlp.OW2, mlp.OB2),
                           (mlp.OB2, mlp.OW3), (mlp.OW3, mlp.OB3))] + [np.array([mlp.OB3])])
    eps = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += eps
        pm[k] -= eps
        fd = (mlp.loss_grad64(pp, x, t)[0] - mlp.loss_grad64(pm, x, t)[0]) / (2 * eps)
        assert fd == pytest.approx(g[k], rel=1e-6, abs=1e-9), k


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -7, -2.5, 1.0 + 2.0 ** -9], np.float32)
    want = np.array([1.0, 1.0, 1.0 + 2.0 ** -6, 1.0 + 2.0 ** -7, -2.5, 1.0], np.float32)
    assert np.array_equal(mlp.bf16(x), want)


def test_bf16_path_close_to_float64():
    rng = np.random.default_rng(4)
    p = mlp.glorot(11)
    x, t, _ = mlp.samples(5, 0, 256)
    l64, g64 = mlp.loss_grad64(p, x, t)
    l16, g16 = mlp.loss_grad_bf16(p, x, t)
    assert l16 == pytest.approx(l64, rel=2e-2)
    assert np.linalg.norm(g16 - g64) <= 5e-2 * np.linalg.norm(g64)
    del rng


def test_labels_closed_forms():
    x, lab, st = mlp.samples(7, 0, 4000)
    i = np.arange(4000, dtype=np.uint64)
    b = 2.0 + 62.0 * mlp.uniform(7, i, 1)
    circle = mlp.uniform(7, i, 0) < 0.5
    sign = np.where(mlp.uniform(7, i, 11) < 0.5, -1.0, 1.0)
    assert np.allclose(lab[circle], (sign / b)[circle].astype(np.float32), rtol=1e-6)
    # ellipse curvature formula against the parametric curve
    a_, b_ = 6.0, 2.5
    t = np.linspace(0.0, 2 * np.pi, 50)
    h = 1e-4
    xp = lambda s: a_ * np.cos(s)  # noqa: E731
    yp = lambda s: b_ * np.sin(s)  # noqa: E731
    x1, y1 = (xp(t + h) - xp(t - h)) / (2 * h), (yp(t + h) - yp(t - h)) / (2 * h)
    x2, y2 = (xp(t + h) - 2 * xp(t) + xp(t - h)) / h ** 2, (yp(t + h) - 2 * yp(t) + yp(t - h)) / h ** 2
    kfd = np.abs(x1 * y2 - y1 * x2) / (x1 ** 2 + y1 ** 2) ** 1.5
    kform = a_ * b_ / (a_ ** 2 * np.sin(t) ** 2 + b_ ** 2 * np.cos(t) ** 2) ** 1.5
    assert np.allclose(kform, kfd, rtol=1e-6)
    # features are the BF16 differences to the centre and the centre - 1/2
    f = mlp.bf16(mlp.features(st))
    assert np.array_equal(f, x)


def test_sharp_stencils_hold_area_fractions():
    seed, n = 9, 64
    _, _, st = mlp.samples(seed, 0, n)
    i = np.arange(n, dtype=np.uint64)
    u = [mlp.uniform(seed, i, k) for k in range(13)]
    ellipse = u[0] >= 0.5
    sharp = ellipse | (u[7] >= 0.5)
    b = 2.0 + 62.0 * u[1]
    a = np.where(ellipse, 1.0 + 2.0 * u[2], 1.0) * b
    phi, t = 2 * np.pi * u[3], 2 * np.pi * u[4]
    px = np.cos(phi) * a * np.cos(t) - np.sin(phi) * b * np.sin(t)
    py = np.sin(phi) * a * np.cos(t) + np.cos(phi) * b * np.sin(t)
    ic, jc = np.floor(px - u[5]), np.floor(py - u[6])
    mm = ((u[12] - 0.5) * 16.0).astype(np.int64)
    sh = np.where(mm < 4, mm, mm + 1)
    shift = u[12] >= 0.5
    ic = np.where(shift, ic + sh % 3 - 1, ic)
    jc = np.where(shift, jc + sh // 3 - 1, jc)
    comp = u[11] < 0.5
    m = 256  # sub-samples per cell width of the window (16x finer)
    s = (np.arange(5 * m) + 0.5) / m - 2.0
    checked = 0
    for k in np.nonzero(sharp)[0]:
        X, Y = np.meshgrid(ic[k] + u[5][k] + s, jc[k] + u[6][k] + s)
        xe = np.cos(phi[k]) * X + np.sin(phi[k]) * Y
        ye = -np.sin(phi[k]) * X + np.cos(phi[k]) * Y
        area = np.count_nonzero((xe / a[k]) ** 2 + (ye / b[k]) ** 2 < 1.0) / m ** 2
        if comp[k]:
            area = 25.0 - area
        assert st[k].sum() == pytest.approx(area, abs=0.05), k
        if not shift[k]:
            assert 0.0 < st[k][12] < 1.0
        checked += 1
    assert checked > 20


def test_diffuse_stencils_centred_in_the_profile():
    seed, n = 13, 2000
    _, _, st = mlp.samples(seed, 0, n)
    i = np.arange(n, dtype=np.uint64)
    diffuse = (mlp.uniform(seed, i, 0) < 0.5) & (mlp.uniform(seed, i, 7) < 0.5)
    c = st[diffuse, 12]
    c = np.minimum(c, 1.0 - c)  # complemented samples mirror the profile
    c = np.concatenate([c, 1.0 - c])
    # the offset point is within +-2/s cells of the interface; the centre of
    # its cell is within half a diagonal of it
    assert c.min() > 0.002 and c.max() < 0.998
    assert np.mean((c >= 0.02) & (c <= 0.98)) > 0.9


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
