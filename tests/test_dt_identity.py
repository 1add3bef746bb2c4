"""The identity the dt reduction's kernel rests on, on the CPU.

decide_dt's advective limit is the minimum over faces of h / (|u| + 1e-8)
(src/decide_dt.c:34-56).  The kernel (csrc/step.hip, k_dt_partial) keeps
each column's largest |u| + 1e-8 and divides once: for h > 0 a correctly
rounded quotient never grows with its divisor, so the smallest quotient is
the quotient by the largest divisor, bit for bit.  fmax passes over a NaN
as fmin does, so a NaN velocity is skipped either way.  This pins that
identity with IEEE float64 division (numpy's, correctly rounded, as the
device's is: tests/test_gpu_parity.py::test_sqrt_div_correctly_rounded).
"""
import numpy as np


def per_face(h, u):
    m = 1.0
    for x in u:
        m = np.fmin(m, h / (abs(x) + 1e-8))
    return m


def per_column(h, u):
    d = 0.0
    for x in u:
        d = np.fmax(d, abs(x) + 1e-8)
    # all NaN: d stays 0 and h / 0 = inf leaves the minimum at 1, as the
    # kernel's division does
    with np.errstate(divide="ignore"):
        return np.fmin(1.0, h / d)


def test_min_of_quotients_is_quotient_by_max_bitwise():
    rng = np.random.default_rng(11)
    for trial in range(2000):
        n = int(rng.integers(1, 40))
        h = float(rng.uniform(1e-6, 1.0)) * 10.0 ** int(rng.integers(-3, 1))
        kind = trial % 4
        if kind == 0:  # wide range
            u = rng.standard_normal(n) * 10.0 ** rng.uniform(-12, 4, n)
        elif kind == 1:  # values a few ulp apart, where the rounding of h / x could tie
            base = float(rng.uniform(0.1, 10.0))
            u = base + np.arange(n) * np.spacing(base) * rng.integers(-3, 4, n)
        elif kind == 2:  # zeros and values near the 1e-8 guard
            u = rng.choice([0.0, -0.0, 1e-8, 3e-9, 1e-12, 2e-8], n)
        else:  # NaNs among finite values
            u = rng.standard_normal(n)
            u[rng.random(n) < 0.3] = np.nan
        assert per_face(h, u) == per_column(h, u), (trial, h, u)
