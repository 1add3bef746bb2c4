"""The learned curvature on the GPU, against the CPU restatement
(oracle/mlp.py) and against the plain curvature it replaces.

* The device rasterises the same training stencils as the restatement.
* The gradients of the BF16-MFMA training kernels equal the restatement's
  with the same BF16 roundings (relative 1e-4 per parameter block).
* Training on 2^21 stencils lowers the loss, and on held-out sharp stencils
  the network is far more accurate than the plain curvature.
* In the solver: on a sharp static drop the network removes most of the
  force imbalance of the first step; on the wide tanh profile of the
  initial condition it leaves the plain curvature in place (the result is
  bitwise the plain one).
"""
import numpy as np
import pytest

from oracle import mlp as omlp
from simplebubblyflowsolver_amd import configs, ic, solver

pytestmark = pytest.mark.gpu

TRAIN = dict(seed=2, nsamples=1 << 21, epochs=40, batch=4096, lr=1e-3)


def test_samples_match_restatement(product):
    net = solver.CurvatureNet(product)
    x, t, st = net.samples(21, 1000, 512)
    xo, to, sto = omlp.samples(21, 1000, 512)
    assert np.array_equal(t, to)
    assert np.abs(st - sto).max() < 1e-6
    assert np.array_equal(x, xo)


def test_gradient_matches_restatement(product):
    net = solver.CurvatureNet(product)
    rng = np.random.default_rng(1)
    p = omlp.glorot(5)
    p[omlp.OB1:omlp.OW2] = 0.1 * rng.standard_normal(64)
    p[omlp.OB2:omlp.OW3] = 0.1 * rng.standard_normal(64)
    p[omlp.OB3] = 0.05
    net.set_params(p)
    x, t, _ = omlp.samples(8, 0, 512)
    loss, g = net.gradient(x, t)
    lo, go = omlp.loss_grad_bf16(p, x, t)
    print(f"loss device {loss:.9e} restatement {lo:.9e}", flush=True)
    assert loss == pytest.approx(lo, rel=1e-5)
    for name, a, b in (("W1", omlp.OW1, omlp.OB1), ("b1", omlp.OB1, omlp.OW2), ("W2", omlp.OW2, omlp.OB2),
                       ("b2", omlp.OB2, omlp.OW3), ("w3", omlp.OW3, omlp.OB3), ("b3", omlp.OB3, omlp.NP)):
        err = np.linalg.norm(g[a:b] - go[a:b]) / np.linalg.norm(go[a:b])
        print(f"{name}: relative gradient difference {err:.2e}", flush=True)
        assert err < 1e-4, name


@pytest.fixture(scope="module")
def trained(product):
    net = solver.CurvatureNet(product)
    hist = net.train(**TRAIN)
    print("training loss per epoch", np.array2string(hist[[0, 1, 9, 19, -1]], precision=3), flush=True)
    return net, hist


def test_training_beats_plain_curvature(trained):
    net, hist = trained
    assert hist[-1] < hist[0] / 5
    x, t, st = omlp.samples(99, 0, 16384)   # held out: another seed
    y, plain = net.predict(st), omlp.plain_curvature(st)
    b5 = st.reshape(-1, 5, 5)
    sharp = (b5.min(axis=(1, 2)) < 1e-3) & (b5.max(axis=(1, 2)) > 1 - 1e-3)
    inr = (np.abs(t) >= 1 / 64) & (np.abs(t) <= 0.5)
    for name, m in (("sharp", sharp & inr), ("diffuse", ~sharp & inr)):
        el = np.mean(np.abs(y[m] - t[m]) / np.abs(t[m]))
        ep = np.mean(np.abs(plain[m] - t[m]) / np.abs(t[m]))
        print(f"{name} stencils ({m.sum()}): relative error learned {el:.3f}, plain {ep:.3f}", flush=True)
        if name == "sharp":
            assert el < ep / 5


def sharp_drop_vof(n, r, sub=16):
    """Gas fraction of each cell inside a circle of radius r at (1/2, 1/2),
    by sub x sub point sampling, with the two wall columns of the NPY layout."""
    xc = (np.arange(n) + 0.5) / n
    s = (np.arange(sub) + 0.5) / sub / n - 0.5 / n
    X = xc[None, :, None, None] + s[None, None, None, :]
    Y = xc[:, None, None, None] + s[None, None, :, None]
    return np.pad(((X - 0.5) ** 2 + (Y - 0.5) ** 2 < r * r).mean(axis=(2, 3)), ((0, 0), (1, 1)))


def drop(product, net, learned, vof, steps):
    w = configs.static_drop(256, steps)
    assert product.lib.sbfs_set_precision(1) == 0
    net.use(learned)
    g = solver.Case(product, w.nx, w.ny, w.lx, w.ly, w.Re, w.Fr, w.denr, w.visr, w.We)
    try:
        g.load_initial_condition(vof)
        for _ in range(steps):
            g.integrate()
        return g.get("ux"), g.get("uy"), g.diagnostics()
    finally:
        g.close()
        net.use(False)
        product.lib.sbfs_set_precision(0)


def test_sharp_drop_first_step_imbalance(product, trained):
    net, _ = trained
    vof = sharp_drop_vof(256, 32 / 256)
    umax = {}
    for learned in (False, True):
        ux, uy, d = drop(product, net, learned, vof, 1)
        umax[learned] = max(np.abs(ux).max(), np.abs(uy).max())
    print(f"sharp static drop, max |u| after one step: plain {umax[False]:.3e}, learned {umax[True]:.3e}",
          flush=True)
    assert umax[True] < umax[False] / 3


def test_wide_profile_keeps_plain_curvature(product, trained):
    net, _ = trained
    vof = ic.bubble_vof(256, 256, 1.0, 1.0, 0.5, 0.5, 32 / 256)
    a = drop(product, net, False, vof, 200)
    b = drop(product, net, True, vof, 200)
    print(f"wide-profile static drop, max |u| after 200 steps: {np.abs(a[0]).max():.3e}", flush=True)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_learned_steps_after_a_larger_case(product, trained):
    """A larger case grows the inference's segment buffer; steps of a
    smaller case captured before it must not replay with the old buffer
    (the step graphs are keyed on the buffer), so the small case repeats
    bitwise."""
    net, _ = trained

    def run(n, steps):
        w = configs.static_drop(n, steps, radius_cells=n / 8)
        assert product.lib.sbfs_set_precision(1) == 0
        net.use(True)
        g = solver.Case(product, w.nx, w.ny, w.lx, w.ly, w.Re, w.Fr, w.denr, w.visr, w.We)
        try:
            g.load_initial_condition(sharp_drop_vof(n, 1 / 8))
            for _ in range(steps):
                g.integrate()
            return g.get("ux"), g.get("uy"), g.get("vof")
        finally:
            g.close()
            net.use(False)
            product.lib.sbfs_set_precision(0)

    first = run(64, 12)
    run(256, 4)
    again = run(64, 12)
    for a, b in zip(first, again):
        assert np.array_equal(a, b)
