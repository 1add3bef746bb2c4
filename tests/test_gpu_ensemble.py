"""Ensembles of independent cases (sbfs_ensemble_*): every case of a batched
step gives the same bits as the case stepped alone with the same options
(partitioned or sequential tridiagonal solve, FP64 or mixed), over steps
that cover both pointer states, the eager steps, the graph capture and its
replays; the autotuner restores the state it times the candidates on."""
import numpy as np
import pytest

from simplebubblyflowsolver_amd import configs, solver

pytestmark = pytest.mark.gpu

FIELDS = ("ux", "uy", "p", "psi0", "vof")


def specs_of(ws):
    return [dict(nx=w.nx, ny=w.ny, lx=w.lx, ly=w.ly, Re=w.Re, Fr=w.Fr, denr=w.denr, visr=w.visr, We=w.We)
            for w in ws]


def small_ensemble():
    ws = configs.ensemble(6, seed=11)
    # sizes from the ensemble's set, plus a 32 x 64 case (one strip, few bands)
    ws[0].nx, ws[0].ny = 32, 64
    return ws


def alone(product, w, steps):
    g = solver.Case(product, w.nx, w.ny, w.lx, w.ly, w.Re, w.Fr, w.denr, w.visr, w.We)
    try:
        g.load_initial_condition(w.vof())
        dts = [g.integrate() for _ in range(steps)]
        return dts, {n: g.get(n) for n in FIELDS}
    finally:
        g.close()


@pytest.mark.parametrize("precision,tdm", [(0, 1), (1, 1), (1, 0)])
def test_ensemble_matches_cases_alone(product, precision, tdm):
    ws = small_ensemble()
    steps = 6
    assert product.lib.sbfs_set_precision(precision) == 0
    assert product.lib.sbfs_set_tdm(tdm, 32) == 0
    try:
        e = solver.Ensemble(product, specs_of(ws))
        try:
            for c, w in zip(e.cases, ws):
                c.load_initial_condition(w.vof())
            dts = np.array([e.step() for _ in range(steps)])
            got = [{n: c.get(n) for n in FIELDS} for c in e.cases]
        finally:
            e.close()
        for k, w in enumerate(ws):
            print(f"case {k}: {w.nx}x{w.ny} Re {w.Re:.1f} We {w.We:.1f} denr {w.denr:.2e} visr {w.visr:.2e}",
                  flush=True)
            dt1, f1 = alone(product, w, steps)
            assert np.array_equal(dts[:, k], np.array(dt1)), k
            for n in FIELDS:
                assert np.array_equal(got[k][n], f1[n]), (k, n)
    finally:
        product.lib.sbfs_set_precision(0)
        product.lib.sbfs_set_tdm(1, 32)


def test_autotune_restores_state(product):
    ws = small_ensemble()
    e = solver.Ensemble(product, specs_of(ws))
    try:
        for c, w in zip(e.cases, ws):
            c.load_initial_condition(w.vof())
        for _ in range(2):
            e.step()
        before = [{n: c.get(n) for n in FIELDS} for c in e.cases]
        band = e.autotune()
        print(f"autotuned band: {band} rows", flush=True)
        assert band in (8, 16, 32, 64, 128, 256)
        after = [{n: c.get(n) for n in FIELDS} for c in e.cases]
        for k in range(len(ws)):
            for n in FIELDS:
                assert np.array_equal(before[k][n], after[k][n]), (k, n)
        for _ in range(2):
            e.step()
        got = [{n: c.get(n) for n in FIELDS} for c in e.cases]
    finally:
        e.close()
    for k, w in enumerate(ws):
        _, f1 = alone(product, w, 4)
        for n in FIELDS:
            assert np.array_equal(got[k][n], f1[n]), (k, n)


def test_ensemble_with_a_wide_case(product):
    """A 4096-column case (its row transforms take two wavefronts per row)
    beside narrow ones: the batched transforms then run 128-thread
    workgroups, two one-wavefront units per workgroup for the narrow cases.
    Every case still gives the bits it gives alone."""
    ws = configs.ensemble(3, seed=5)
    ws[0].nx, ws[0].ny, ws[0].lx, ws[0].ly = 4096, 16, 1.0, 16 / 4096
    ws[1].nx, ws[1].ny = 64, 128
    ws[2].nx, ws[2].ny = 128, 256
    ws[0].bubbles = [(0.5, 0.5 * 16 / 4096, 0.001)]
    steps = 4
    e = solver.Ensemble(product, specs_of(ws))
    try:
        for c, w in zip(e.cases, ws):
            c.load_initial_condition(w.vof())
        dts = np.array([e.step() for _ in range(steps)])
        got = [{n: c.get(n) for n in FIELDS} for c in e.cases]
    finally:
        e.close()
    for k, w in enumerate(ws):
        dt1, f1 = alone(product, w, steps)
        assert np.array_equal(dts[:, k], np.array(dt1)), k
        for n in FIELDS:
            assert np.array_equal(got[k][n], f1[n]), (k, n)
