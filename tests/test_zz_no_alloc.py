"""decide_dt_noalloc (include/sbfs.h): the time step from the library's
fixed device array, with no device allocation, against decide_dt on the
same state, bitwise, on two calls in a row with the second on a smaller
case; and its refusal of a domain without a device context."""
import ctypes as C

import numpy as np
import pytest

from harness import initial_state
from simplebubblyflowsolver_amd import abi, solver

pytestmark = pytest.mark.gpu


def bind(lib):
    f = lib.decide_dt_noalloc
    f.argtypes = [C.POINTER(abi.DomainT), C.POINTER(abi.FluidT), C.POINTER(abi.InterfaceT),
                  C.POINTER(C.c_double)]
    f.restype = C.c_int
    return f


def noalloc_dt(product, case) -> float:
    dt = C.c_double(0.0)
    rc = bind(product.lib)(C.byref(case.domain), C.byref(case.fluid), C.byref(case.interface), C.byref(dt))
    product.check(rc, "decide_dt_noalloc")
    return dt.value


@pytest.mark.parametrize("velocity", [0.0, 2.0])
def test_noalloc_dt_matches_decide_dt(product, velocity):
    sizes = [(512, 1024, 3), (64, 128, 4)]  # the second call on a smaller case
    for nx, ny, seed in sizes:
        print(f"decide_dt_noalloc: {nx} x {ny}, seed {seed}, velocity {velocity}", flush=True)
        c = solver.Case(product, nx, ny, 1.0, 2.0, 35.0, 1.0, 1e-3, 1e-2, 10.0)
        try:
            initial_state(c, seed=seed, velocity=velocity)
            want = c.decide_dt()
            got = noalloc_dt(product, c)
            assert np.float64(got).tobytes() == np.float64(want).tobytes(), (nx, ny, got, want)
            # and again after a step moved the fields
            c.integrate()
            got = noalloc_dt(product, c)
            want = c.decide_dt()
            assert np.float64(got).tobytes() == np.float64(want).tobytes(), (nx, ny, got, want)
        finally:
            c.close()


def test_noalloc_dt_refuses_a_domain_without_device_context(product):
    c = solver.Case(product, 64, 128, 1.0, 2.0, 35.0, 1.0, 1e-3, 1e-2, 10.0)
    try:
        other = abi.DomainT()
        C.memmove(C.byref(other), C.byref(c.domain), C.sizeof(abi.DomainT))
        dt = C.c_double(-1.0)
        rc = bind(product.lib)(C.byref(other), C.byref(c.fluid), C.byref(c.interface), C.byref(dt))
        assert rc != 0 and dt.value == -1.0
        assert "no device context" in product.error_text()
    finally:
        c.close()
