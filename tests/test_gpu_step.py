"""Whole time steps on the GPU against the CPU oracle, and the step's
invariants at full size."""
import ctypes as C
import os
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

from harness import copy_state, initial_state, new_pair
from simplebubblyflowsolver_amd import configs, solver

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def demean(a):
    return a - a.mean()


def set_path(product, path):
    assert product.lib.sbfs_set_path(path) == 0


def set_tdm(product, kind, chunk=32):
    product.lib.sbfs_set_tdm.argtypes = [C.c_int, C.c_int]
    assert product.lib.sbfs_set_tdm(kind, chunk) == 0


@pytest.mark.parametrize("path", [0, 1])
def test_rising_bubble_steps_match_oracle(oracle_mod, product, path):
    """64 x 128 rising bubble from rest, 20 steps: velocity and vof within
    1e-10 relative of the oracle; p and psi compared after removing their
    means (the pressure gauge)."""
    w = configs.rising_bubble(64, 128)
    oracle_mod.set_poisson(dct=oracle_mod.DCT_FFT, tdm=oracle_mod.TDM_SEQUENTIAL)
    set_path(product, path)
    o, g = new_pair(oracle_mod, product, w.nx, w.ny)
    try:
        o.load_initial_condition(w.vof())
        copy_state(o, g)
        for s in range(20):
            dto = o.integrate()
            dtg = g.integrate()
            assert dtg == pytest.approx(dto, rel=1e-12)
        for n in ("ux", "uy", "vof"):
            assert rel(g.get(n), o.get(n)) <= 1e-10, n
        for n in ("p", "psi0"):
            assert rel(demean(g.get(n)), demean(o.get(n))) <= 1e-10, n
        do, dg = o.diagnostics(), g.diagnostics()
        assert dg["volume"] == pytest.approx(do["volume"], rel=1e-13)
        assert dg["rise_velocity"] == pytest.approx(do["rise_velocity"], rel=1e-9)
        assert dg["divmax"] < 1e-10
    finally:
        o.close()
        g.close()
        set_path(product, 1)


def test_swarm_invariants(product):
    """Bubble swarm at 512 x 1024 (16 bubbles), 10 steps on the GPU only:
    gas volume conserved to 1e-12 relative, max |div u| at round-off, vof
    bounded."""
    w = configs.bubble_swarm(512, 1024, 16, steps=10, seed=5)
    g = solver.Case(product, w.nx, w.ny, w.lx, w.ly, w.Re, w.Fr, w.denr, w.visr, w.We)
    try:
        g.load_initial_condition(w.vof())
        d0 = g.diagnostics()
        for _ in range(w.steps):
            g.integrate()
        d = g.diagnostics()
        assert abs(d["volume"] - d0["volume"]) <= 1e-12 * d0["volume"]
        assert d["divmax"] < 1e-9
        assert d["vofmin"] > -1e-6 and d["vofmax"] < 1 + 1e-6
    finally:
        g.close()


@pytest.mark.parametrize("nx,ny", [(64, 128), (32, 66)])
def test_fused_path_matches_staged_bitwise(oracle_mod, product, nx, ny):
    """The fused kernels recompute every intermediate with the staged
    kernels' device functions and operation order: 8 steps from a state
    with random velocities agree bitwise in every state field (ny / nx not
    multiples of the tile size exercise the ragged tiles).  The fused case
    runs its first two steps eagerly, captures the two pointer states into
    graphs on steps 3-4 and replays them on steps 5-8, with the staged
    case's eager work and synchronisations in between."""
    from harness import initial_state
    o, _ = new_pair(oracle_mod, product, nx, ny)
    a = solver.Case(product, nx, ny, 1.0, 2.0, 35.0, 1.0, 1e-3, 1e-2, 10.0)
    b = solver.Case(product, nx, ny, 1.0, 2.0, 35.0, 1.0, 1e-3, 1e-2, 10.0)
    try:
        initial_state(o, 7, velocity=0.05)
        copy_state(o, a)
        copy_state(o, b)
        for _ in range(8):
            set_path(product, 0)
            dta = a.integrate()
            set_path(product, 1)
            dtb = b.integrate()
            assert dta == dtb
        for n in ("vof", "ux", "uy", "p", "psi0", "den1", "txy"):
            x, y = a.get(n), b.get(n)
            assert np.array_equal(x, y), (n, np.abs(x - y).max())
    finally:
        o.close()
        a.close()
        b.close()
        set_path(product, 1)


UNFOLDED = """
import sys
import numpy as np
sys.path.insert(0, sys.argv[1])
sys.path.insert(0, sys.argv[1] + "/tests")
from harness import copy_state, initial_state
from simplebubblyflowsolver_amd import solver
be = solver.product_backend()
be.set_dt_coefs(0.2, 0.5, 0.95)
a = solver.Case(be, 64, 128, 1.0, 2.0, 35.0, 1.0, 1e-3, 1e-2, 10.0)
b = solver.Case(be, 64, 128, 1.0, 2.0, 35.0, 1.0, 1e-3, 1e-2, 10.0)
initial_state(a, 7, velocity=0.05)
copy_state(a, b)
for _ in range(8):
    assert be.lib.sbfs_set_path(0) == 0
    dta = a.integrate()
    assert be.lib.sbfs_set_path(1) == 0
    dtb = b.integrate()
    assert dta == dtb
for n in ("vof", "ux", "uy", "p", "psi0", "den1", "txy"):
    x, y = a.get(n), b.get(n)
    assert np.array_equal(x, y), (n, np.abs(x - y).max())
print("unfolded fused path: 8 steps bitwise equal to the staged path", flush=True)
"""


def test_unfolded_fused_path_matches_staged_bitwise():
    """With SBFS_MARCH_FOLD=0 the fused step runs k_post after every stage
    instead of folding each projection into the next stage kernel, and no
    stage kernel writes a pressure: 8 steps bitwise equal to the staged
    path.  The switch is read once per process, so the comparison runs in a
    child process."""
    env = dict(os.environ, SBFS_MARCH_FOLD="0")
    root = str(Path(__file__).resolve().parents[1])
    r = subprocess.run([sys.executable, "-c", UNFOLDED, root], env=env, capture_output=True, text=True,
                       timeout=240)
    print(r.stdout.strip())
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.parametrize("nx,ny,bubbles,steps,bands", [
    (512, 1024, 16, 8, ("11", "7")),
    # band 4: 36 x 1024 units, more than the sort takes in one pass (8192)
    (2048, 4096, 64, 3, ("", "4")),
])
def test_dispatch_order_does_not_change_results(product, nx, ny, bubbles, steps, bands):
    """The stage kernels' units run heaviest first (each step ordered by the
    interface rows the previous step's stage kernel counted) or in the
    grid's order, with bands of different height: a swarm stepped both ways
    agrees bitwise in every state field (the first case in grid order with
    the first band height, "" for the default, the second ordered)."""
    import os
    w = configs.bubble_swarm(nx, ny, bubbles, steps=steps, seed=3)
    print(f"{w.name}: {steps} steps, bands {bands}", flush=True)
    cases = [solver.Case(product, w.nx, w.ny, w.lx, w.ly, w.Re, w.Fr, w.denr, w.visr, w.We) for _ in range(2)]
    saved = {k: os.environ.get(k) for k in ("SBFS_MARCH_ORDER", "SBFS_MARCH_BAND")}
    try:
        for g in cases:
            g.load_initial_condition(w.vof())
        for _ in range(w.steps):
            dts = []
            for g, order, band in zip(cases, ("0", "1"), bands):
                os.environ["SBFS_MARCH_ORDER"] = order
                if band:
                    os.environ["SBFS_MARCH_BAND"] = band
                else:
                    os.environ.pop("SBFS_MARCH_BAND", None)
                dts.append(g.integrate())
            assert dts[0] == dts[1]
        for n in ("vof", "ux", "uy", "p", "psi0", "den1", "txy"):
            x, y = cases[0].get(n), cases[1].get(n)
            assert np.array_equal(x, y), (n, np.abs(x - y).max())
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
        for g in cases:
            g.close()


@pytest.mark.parametrize("mixed", [0, 1])
def test_streaming_variants_match_ordinary_bitwise(product, mixed):
    """Large grids take the kernels' streaming variants (non-temporal loads and
    stores of the rows a kernel touches once); SBFS_STREAM forces either
    choice.  A 1024 x 2048 swarm stepped with the streaming variants and with
    ordinary accesses agrees bitwise in every state field, in FP64 and in
    mixed precision."""
    import os
    w = configs.bubble_swarm(1024, 2048, 16, steps=4, seed=5)
    print(f"{w.name}: {w.steps} steps, mixed={mixed}", flush=True)
    cases = [solver.Case(product, w.nx, w.ny, w.lx, w.ly, w.Re, w.Fr, w.denr, w.visr, w.We) for _ in range(2)]
    saved = os.environ.get("SBFS_STREAM")
    assert product.lib.sbfs_set_precision(mixed) == 0
    try:
        for g in cases:
            g.load_initial_condition(w.vof())
        for _ in range(w.steps):
            dts = []
            for g, flag in zip(cases, ("1", "0")):
                os.environ["SBFS_STREAM"] = flag
                dts.append(g.integrate())
            assert dts[0] == dts[1]
        for n in ("vof", "ux", "uy", "p", "psi0", "den1", "txy"):
            x, y = cases[0].get(n), cases[1].get(n)
            assert np.array_equal(x, y), (n, np.abs(x - y).max())
    finally:
        assert product.lib.sbfs_set_precision(0) == 0
        if saved is None:
            os.environ.pop("SBFS_STREAM", None)
        else:
            os.environ["SBFS_STREAM"] = saved
        for g in cases:
            g.close()


def test_repo_default_rising_bubble_100_steps(oracle_mod, product):
    """The 128 x 256 rising bubble (the reference's CPU-runnable case), 100
    steps on the default (fused) path against the oracle: velocity and vof
    within 1e-10 relative, p and psi after removing their means, the
    rise-velocity history overlaid to 1e-9 and the volume to 1e-13."""
    w = configs.rising_bubble(128, 256)
    oracle_mod.set_poisson(dct=oracle_mod.DCT_FFT, tdm=oracle_mod.TDM_SEQUENTIAL)
    o, g = new_pair(oracle_mod, product, w.nx, w.ny)
    try:
        o.load_initial_condition(w.vof())
        copy_state(o, g)
        v0 = o.diagnostics()["volume"]
        for s in range(100):
            dto, dtg = o.integrate(), g.integrate()
            assert dtg == pytest.approx(dto, rel=1e-12)
            if s % 20 == 19:
                do, dg = o.diagnostics(), g.diagnostics()
                assert dg["rise_velocity"] == pytest.approx(do["rise_velocity"], rel=1e-9)
                assert abs(dg["volume"] - v0) <= 1e-13 * v0
        for n in ("ux", "uy", "vof"):
            assert rel(g.get(n), o.get(n)) <= 1e-10, n
        for n in ("p", "psi0"):
            assert rel(demean(g.get(n)), demean(o.get(n))) <= 1e-10, n
    finally:
        o.close()
        g.close()


def test_repo_default_rising_bubble_2000_steps(oracle_mod, product):
    """The reference's own case in full: the 128 x 256 rising bubble for
    2000 steps on the default (fused, partitioned) path against the oracle
    in the reference's solve order.  The rise-velocity and centroid
    histories (every 50 steps) overlay, the gas volume stays at its initial
    value, and the final fields agree.  Measured on MI355X: history 1.5e-14,
    centroid 5e-15, ux / uy / p 2.4e-14, vof 3e-16, psi 6e-12 (the
    partitioned against the sequential tridiagonal order); the bounds leave
    a margin of about 50x."""
    w = configs.rising_bubble(128, 256)
    oracle_mod.set_poisson(dct=oracle_mod.DCT_FFT, tdm=oracle_mod.TDM_SEQUENTIAL)
    o, g = new_pair(oracle_mod, product, w.nx, w.ny)
    try:
        o.load_initial_condition(w.vof())
        copy_state(o, g)
        v0 = o.diagnostics()["volume"]
        ur_o, ur_g, xc_o, xc_g = [], [], [], []
        for s in range(2000):
            o.integrate()
            g.integrate()
            if s % 50 == 49:
                do, dg = o.diagnostics(), g.diagnostics()
                ur_o.append(do["rise_velocity"])
                ur_g.append(dg["rise_velocity"])
                xc_o.append(do["centroid"])
                xc_g.append(dg["centroid"])
                assert abs(dg["volume"] - v0) <= 1e-12 * v0
        ur_o, ur_g = np.array(ur_o), np.array(ur_g)
        e_ur = np.abs(ur_g - ur_o).max() / np.abs(ur_o).max()
        e_xc = np.abs(np.array(xc_g) - np.array(xc_o)).max()
        errs = {n: rel(g.get(n), o.get(n)) for n in ("ux", "uy", "vof")}
        errs.update({n: rel(demean(g.get(n)), demean(o.get(n))) for n in ("p", "psi0")})
        print(f"2000 steps: rise velocity {ur_o[-1]:.6e}, history error {e_ur:.2e}, centroid error {e_xc:.2e}, "
              f"fields {errs}", flush=True)
        assert e_ur <= 1e-12 and e_xc <= 1e-12
        for n, e in errs.items():
            assert e <= (3e-10 if n == "psi0" else 1e-12), n
    finally:
        o.close()
        g.close()


@pytest.mark.parametrize("which", ["swarm", "large"])
def test_full_size_swarm_invariants(product, which):
    """The benchmark swarm (2048 x 4096, 64 bubbles) and the large swarm
    (4096 x 8192, 1024 bubbles), 5 steps at full size: gas volume conserved
    to 1e-12 relative, max |div u| at round-off, vof within [0, 1] to the
    band tolerance."""
    w = configs.bubble_swarm(2048, 4096, 64, seed=1) if which == "swarm" else configs.large_swarm()
    print(f"{w.name}: {w.nx} x {w.ny}", flush=True)
    g = solver.Case(product, w.nx, w.ny, w.lx, w.ly, w.Re, w.Fr, w.denr, w.visr, w.We)
    try:
        g.load_initial_condition(w.vof())
        d0 = g.diagnostics()
        for _ in range(5):
            g.integrate()
        d = g.diagnostics()
        assert abs(d["volume"] - d0["volume"]) <= 1e-12 * d0["volume"]
        assert d["divmax"] < 1e-9
        assert d["vofmin"] > -1e-6 and d["vofmax"] < 1 + 1e-6
    finally:
        g.close()


def test_large_swarm_matches_oracle(oracle_mod, product):
    """4096 x 8192 with 1024 bubbles (4096-point DCT rows, 64-row chunks of
    the partitioned solve), 2 steps on the GPU against the oracle on the
    host: velocity and vof within 1e-10 relative, p and psi after removing
    their means."""
    w = configs.large_swarm()
    print(f"{w.name}: oracle and GPU, 2 steps", flush=True)
    oracle_mod.set_threads(16)
    oracle_mod.set_poisson(dct=oracle_mod.DCT_FFT, tdm=oracle_mod.TDM_SEQUENTIAL)
    o, g = new_pair(oracle_mod, product, w.nx, w.ny, phys=dict(Re=w.Re, We=w.We))
    try:
        vof = w.vof()
        o.load_initial_condition(vof)
        g.load_initial_condition(vof)
        for _ in range(2):
            assert g.integrate() == pytest.approx(o.integrate(), rel=1e-12)
        for n in ("ux", "uy", "vof"):
            assert rel(g.get(n), o.get(n)) <= 1e-10, n
        for n in ("p", "psi0"):
            assert rel(demean(g.get(n)), demean(o.get(n))) <= 1e-10, n
    finally:
        o.close()
        g.close()
