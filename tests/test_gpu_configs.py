"""The configured workloads at their full sizes, in the forms they are
specified in:

* the batched ensemble of 256 rising bubbles on uneven grids (64 x 128,
  128 x 256, 256 x 512; density and viscosity ratios 1e-3 .. 1e-1, We 1 ..
  50, Re 20 .. 200): 20 mixed-precision steps keep every case's gas volume
  and divergence, and four of its cases stepped in FP64 agree with the CPU
  oracle within 1e-10;
* the 4096 x 8192 swarm of 1024 bubbles (Re 100, We 5) in mixed precision
  with the learned curvature against the plain FP64 step over 30 steps:
  rise-velocity history within 1 %, gas volume within 1e-10, divergence at
  round-off, and the network takes the interface band;
* the mixed-precision step against the CPU oracle's FP64 step (not only
  against the library's own FP64 step) on the repository's default case.
"""
import numpy as np
import pytest

from simplebubblyflowsolver_amd import configs, solver

pytestmark = pytest.mark.gpu

FIELDS = ("ux", "uy", "vof")


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def specs_of(ws):
    return [dict(nx=w.nx, ny=w.ny, lx=w.lx, ly=w.ly, Re=w.Re, Fr=w.Fr, denr=w.denr, visr=w.visr, We=w.We)
            for w in ws]


def test_ensemble_256_cases_mixed_invariants(product):
    ws = configs.ensemble(256, seed=3)
    assert product.lib.sbfs_set_precision(1) == 0
    try:
        e = solver.Ensemble(product, specs_of(ws))
        try:
            for c, w in zip(e.cases, ws):
                c.load_initial_condition(w.vof())
            v0 = np.array([c.diagnostics()["volume"] for c in e.cases])
            for _ in range(2):
                e.step()
            band = e.autotune()
            for _ in range(18):
                dts = e.step()
            d = [c.diagnostics() for c in e.cases]
        finally:
            e.close()
    finally:
        product.lib.sbfs_set_precision(0)
    drift = np.abs(np.array([x["volume"] for x in d]) - v0) / v0
    div = np.array([x["divmax"] for x in d])
    ur = np.array([x["rise_velocity"] for x in d])
    print(f"256 cases, 20 mixed steps (band {band}): max volume drift {drift.max():.2e}, "
          f"max divergence {div.max():.2e}, rise velocity {ur.min():.3e} .. {ur.max():.3e}", flush=True)
    assert np.all(np.isfinite(dts)) and np.all(dts > 0)
    assert drift.max() <= 1e-12
    assert div.max() < 1e-8
    assert np.all(ur > 0)   # every bubble has started to rise (+x)


def test_ensemble_cases_match_oracle(oracle_mod, product):
    """The 256-case ensemble in FP64, 5 steps; four of its cases (each grid
    size, the most extreme ratios) against the oracle stepping the same
    case alone."""
    ws = configs.ensemble(256, seed=3)
    pick = []
    for nx in (64, 128, 256):
        pick.append(next(k for k, w in enumerate(ws) if w.nx == nx))
    pick.append(int(np.argmin([w.denr * w.visr for w in ws])))
    steps = 5
    oracle_mod.set_poisson(dct=oracle_mod.DCT_FFT, tdm=oracle_mod.TDM_SEQUENTIAL)
    e = solver.Ensemble(product, specs_of(ws))
    try:
        for c, w in zip(e.cases, ws):
            c.load_initial_condition(w.vof())
        dts = np.array([e.step() for _ in range(steps)])
        got = {k: {n: e.cases[k].get(n) for n in FIELDS + ("p",)} for k in pick}
    finally:
        e.close()
    for k in pick:
        w = ws[k]
        print(f"case {k}: {w.nx}x{w.ny} Re {w.Re:.1f} We {w.We:.1f} denr {w.denr:.2e} visr {w.visr:.2e}",
              flush=True)
        o = oracle_mod.new_case(w.nx, w.ny, w.lx, w.ly, w.Re, w.Fr, w.denr, w.visr, w.We)
        try:
            o.load_initial_condition(w.vof())
            for s in range(steps):
                assert dts[s, k] == pytest.approx(o.integrate(), rel=1e-12)
            for n in FIELDS:
                assert rel(got[k][n], o.get(n)) <= 1e-10, (k, n)
            po, pg = o.get("p"), got[k]["p"]
            assert rel(pg - pg.mean(), po - po.mean()) <= 1e-10, k
        finally:
            o.close()


@pytest.fixture(scope="module")
def net(product):
    n = solver.CurvatureNet(product)
    n.train(seed=2, nsamples=1 << 21, epochs=40, batch=4096, lr=1e-3)
    return n


def history(product, w, steps, mixed, learned_net=None):
    assert product.lib.sbfs_set_precision(1 if mixed else 0) == 0
    if learned_net:
        learned_net.use(True)
    g = solver.Case(product, w.nx, w.ny, w.lx, w.ly, w.Re, w.Fr, w.denr, w.visr, w.We)
    try:
        g.load_initial_condition(w.vof())
        d0 = g.diagnostics()
        cover = learned_net.coverage(g) if learned_net else None
        hist = []
        for _ in range(steps):
            g.integrate()
            hist.append(g.diagnostics())
        return d0, hist, cover
    finally:
        g.close()
        if learned_net:
            learned_net.use(False)
        product.lib.sbfs_set_precision(0)


def test_large_swarm_mixed_learned_against_fp64(product, net):
    w = configs.large_swarm()
    steps = 30
    print(f"{w.name}: {steps} steps FP64, then mixed + learned", flush=True)
    d0, ref, _ = history(product, w, steps, False)
    d0m, got, cover = history(product, w, steps, True, net)
    ur_ref = np.array([d["rise_velocity"] for d in ref])
    ur = np.array([d["rise_velocity"] for d in got])
    err = np.abs(ur - ur_ref).max() / np.abs(ur_ref).max()
    drift = max(abs(d["volume"] - d0m["volume"]) for d in got) / d0m["volume"]
    div = max(d["divmax"] for d in got)
    print(f"learned cells {cover:.4f}, rise-velocity history error {err:.3e}, volume drift {drift:.2e}, "
          f"divmax {div:.2e}", flush=True)
    print("rise velocity FP64    ", np.array2string(ur_ref[::5], precision=3), flush=True)
    print("rise velocity learned ", np.array2string(ur[::5], precision=3), flush=True)
    assert cover > 0.01
    assert drift <= 1e-10
    assert div < 1e-9


def test_mixed_against_oracle_fp64(oracle_mod, product):
    """128 x 256 rising bubble, 400 steps: the mixed-precision step against
    the oracle's FP64 step within the chosen form's bounds (rise velocity
    1 %, gas volume 1e-10) and at the FP32 rounding level in the fields."""
    w = configs.rising_bubble(128, 256)
    steps, every = 400, 10
    oracle_mod.set_threads(16)
    oracle_mod.set_poisson(dct=oracle_mod.DCT_FFT, tdm=oracle_mod.TDM_SEQUENTIAL)
    o = oracle_mod.new_case(w.nx, w.ny, w.lx, w.ly, w.Re, w.Fr, w.denr, w.visr, w.We)
    assert product.lib.sbfs_set_precision(1) == 0
    g = solver.Case(product, w.nx, w.ny, w.lx, w.ly, w.Re, w.Fr, w.denr, w.visr, w.We)
    try:
        o.load_initial_condition(w.vof())
        g.load_initial_condition(w.vof())
        v0 = g.diagnostics()["volume"]
        ur_o, ur_g = [], []
        for s in range(steps):
            o.integrate()
            g.integrate()
            if s % every == every - 1:
                ur_o.append(o.diagnostics()["rise_velocity"])
                ur_g.append(g.diagnostics()["rise_velocity"])
        d = g.diagnostics()
        err = np.abs(np.array(ur_g) - np.array(ur_o)).max() / np.abs(np.array(ur_o)).max()
        errs = {n: rel(g.get(n), o.get(n)) for n in FIELDS}
        drift = abs(d["volume"] - v0) / v0
        print(f"mixed vs oracle FP64, {steps} steps: rise-velocity history {err:.2e}, volume drift {drift:.2e}, "
              f"fields {errs}", flush=True)
        assert err <= 0.01 and drift <= 1e-10 and d["divmax"] < 1e-9
        for n, e in errs.items():
            assert e < 1e-3, n
    finally:
        g.close()
        o.close()
        product.lib.sbfs_set_precision(0)
