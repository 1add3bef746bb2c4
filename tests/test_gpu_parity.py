"""GPU parity of every stage of the time step against the CPU oracle.

Stage-isolated: before each stage the oracle's whole state (every array of
fluid_t and interface_t) is copied to the GPU, both run the stage, and every
array is compared.  Stages without transcendental functions must agree
bitwise (same operation order, no FMA contraction, correctly rounded FP64
division and sqrt on both sides).  The THINC intercept (log) and the THINC
fluxes (tanh) use the GPU's and glibc's own libm: they may differ in the
last bits, so those arrays, and the vof update built from the fluxes, are
held to 1e-13 of the array's largest magnitude (element-wise relative
differences are larger where the flux divergence cancels).
"""
import ctypes as C

import numpy as np
import pytest

from harness import (STAGES, compare, copy_state, initial_state, new_pair, run_stage,
                     ulp_diff)

pytestmark = pytest.mark.gpu

# arrays that may carry libm (tanh / log) last-bit differences after a stage
LIBM_ARRAYS = {
    "curvature_tensor": {"normal"},
    "update_vof": {"fluxx", "fluxy", "src_a", "vof"},
}


def set_tdm(product, kind, chunk=32):
    product.lib.sbfs_set_tdm.argtypes = [C.c_int, C.c_int]
    assert product.lib.sbfs_set_tdm(kind, chunk) == 0


@pytest.mark.parametrize("nx,ny,seed", [(32, 64, 0), (64, 32, 1), (16, 128, 2)])
def test_stages_isolated(oracle_mod, product, nx, ny, seed):
    oracle_mod.set_poisson(dct=oracle_mod.DCT_FFT, tdm=oracle_mod.TDM_SEQUENTIAL)
    set_tdm(product, 0)
    o, g = new_pair(oracle_mod, product, nx, ny)
    try:
        initial_state(o, seed)
        copy_state(o, g)
        dt = o.decide_dt()
        assert g.decide_dt() == dt
        failures = []
        for k in range(3):
            for st in STAGES:
                copy_state(o, g)
                run_stage(o, st, k, dt)
                run_stage(g, st, k, dt)
                for name, (dmax, rmax, nd) in compare(o, g).items():
                    if name in LIBM_ARRAYS.get(st, ()):
                        scale = np.abs(o.get(name)).max()
                        if dmax > 1e-13 * scale:
                            failures.append((k, st, name, dmax, rmax, nd))
                    elif dmax != 0.0:
                        failures.append((k, st, name, dmax, rmax, nd))
        assert not failures, failures[:10]
    finally:
        o.close()
        g.close()
        set_tdm(product, 1)


@pytest.mark.parametrize("nx,ny,chunk", [(32, 128, 16), (16, 48, 16), (64, 96, 8), (8, 2048, 16),
                                         (8, 4096, 32), (8, 8192, 64)])
def test_partitioned_tdm_matches_oracle(oracle_mod, product, nx, ny, chunk):
    """GPU partitioned solver vs the oracle's restatement of the same
    algorithm: bitwise; and vs the reference-order solver: 1e-11 relative
    (reduced systems of 8, 3, 12, 128, 128 and 128 separators; the last one
    has the 4096 x 8192 swarm's 8192-row columns and 64-row chunks)."""
    set_tdm(product, 1, chunk)
    oracle_mod.set_poisson(dct=oracle_mod.DCT_FFT, tdm=oracle_mod.TDM_PARTITIONED, chunk=chunk)
    o, g = new_pair(oracle_mod, product, nx, ny)
    try:
        initial_state(o, 3)
        o.compute_density(1)
        copy_state(o, g)
        dt = o.decide_dt()
        dt_old, dt_new = 0.5 * dt, 0.4 * dt
        o.compute_potential(dt_old, dt_new)
        g.compute_potential(dt_old, dt_new)
        a, b = o.get("psi1"), g.get("psi1")
        assert np.array_equal(a, b), np.abs(a - b).max()
        # the reference's sequential order gives the same field to round-off
        oracle_mod.set_poisson(dct=oracle_mod.DCT_FFT, tdm=oracle_mod.TDM_SEQUENTIAL)
        copy_state(g, o)
        o.compute_potential(dt_old, dt_new)
        c = o.get("psi1")
        # both orders round; the lowest modes' conditioning grows like ny^2
        # (tests/test_oracle_pins.py pins each against extended precision)
        assert np.abs(c - b).max() <= 1e-11 * max(1.0, (ny / 2048) ** 2) * np.abs(c).max()
    finally:
        o.close()
        g.close()
        oracle_mod.set_poisson()
        set_tdm(product, 1, 32)


@pytest.mark.parametrize("n", [2, 4, 8, 16, 64, 128, 1024, 2048, 4096])
def test_dct_rows_bitwise(oracle_mod, product, n):
    rng = np.random.default_rng(n)
    x = rng.normal(size=(5, n))
    f = product.lib.sbfs_dct_rows
    f.argtypes = [C.c_int, C.c_size_t, C.c_size_t, C.c_void_p, C.c_void_p]
    for inverse, ref in ((0, oracle_mod.dct2), (1, oracle_mod.dct3)):
        out = np.empty_like(x)
        assert f(inverse, n, 5, x.ctypes.data, out.ctypes.data) == 0
        expect = ref(x, oracle_mod.DCT_FFT)
        assert np.array_equal(out, expect), (inverse, np.abs(out - expect).max())


def test_sqrt_div_correctly_rounded(oracle_mod, product):
    """dvof normalisation needs correctly rounded sqrt and division: random
    vof fields must give bitwise identical corner gradients."""
    nx, ny = 64, 64
    o, g = new_pair(oracle_mod, product, nx, ny)
    try:
        rng = np.random.default_rng(11)
        vof = rng.uniform(0, 1, size=(ny, nx + 2))
        o.load_initial_condition(vof)
        copy_state(o, g)
        o.compute_curvature_tensor()
        g.compute_curvature_tensor()
        assert np.array_equal(o.get("dvof"), g.get("dvof"))
        assert np.array_equal(o.get("curv"), g.get("curv"))
        assert ulp_diff(o.get("normal"), g.get("normal")) <= 4
    finally:
        o.close()
        g.close()
