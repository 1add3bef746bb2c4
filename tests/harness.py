"""Shared set-up for the parity tests: identical states on the oracle and
on the HIP library, stage-by-stage comparison helpers."""
from __future__ import annotations

import numpy as np

from simplebubblyflowsolver_amd import abi, ic, solver

FIELD_NAMES = list(abi.FIELDS.keys())

PHYS = dict(Re=35.0, Fr=1.0, denr=1e-3, visr=1e-2, We=10.0)


def smooth_field(shape, rng, amp=1.0, modes=3):
    """Sum of a few random sinusoids over a (rows, cols) grid."""
    r, c = shape[:2]
    y = np.linspace(0.0, 2 * np.pi, r, endpoint=False)[:, None]
    x = np.linspace(0.0, np.pi, c)[None, :]
    out = np.zeros((r, c))
    for _ in range(modes):
        ky, kx = rng.integers(1, 4), rng.integers(1, 4)
        out += rng.normal() * np.sin(ky * y + rng.uniform(0, 6.28)) * np.cos(kx * x + rng.uniform(0, 6.28))
    return amp * out


def initial_state(case: solver.Case, seed: int = 0, velocity: float = 0.3,
                  cx=0.5, cy=0.5, r=0.25, lx=1.0, ly=2.0):
    """Bubble vof plus smooth random velocity / pressure / potential."""
    rng = np.random.default_rng(seed)
    nx, ny = case.nx, case.ny
    vof = ic.bubble_vof(nx, ny, lx, ly, cx, cy, r)
    ux = smooth_field((ny, nx + 1), rng, velocity)
    uy = smooth_field((ny, nx + 2), rng, velocity)
    p = smooth_field((ny, nx + 2), rng, 0.5)
    psi = smooth_field((ny, nx + 2), rng, 1e-3)
    case.load_initial_condition(vof, ux, uy, p, psi)


def new_pair(oracle_mod, product, nx, ny, lx=1.0, ly=2.0, phys=None):
    ph = dict(PHYS)
    if phys:
        ph.update(phys)
    o = oracle_mod.new_case(nx, ny, lx, ly, **ph)
    g = solver.Case(product, nx, ny, lx, ly, ph["Re"], ph["Fr"], ph["denr"], ph["visr"], ph["We"])
    return o, g


def copy_state(src: solver.Case, dst: solver.Case, names=None):
    for n in names or FIELD_NAMES:
        dst.set(n, src.get(n))


def compare(o: solver.Case, g: solver.Case, names=None, rtol=0.0):
    """{name: (max abs diff, max rel diff, n differing)} for every field."""
    out = {}
    for n in names or FIELD_NAMES:
        a, b = o.get(n), g.get(n)
        d = np.abs(a - b)
        scale = np.maximum(np.abs(a), 1e-300)
        out[n] = (float(d.max()), float((d / scale).max()), int(np.count_nonzero(d > rtol * scale)))
    return out


def ulp_diff(a, b):
    """Largest distance in units in the last place between two arrays."""
    ia = a.view(np.int64).astype(np.int64)
    ib = b.view(np.int64).astype(np.int64)
    ia = np.where(ia < 0, np.int64(-2**63) - ia, ia)
    ib = np.where(ib < 0, np.int64(-2**63) - ib, ib)
    return int(np.max(np.abs(ia - ib))) if a.size else 0


STAGES = [
    "density0", "viscosity", "curvature_tensor", "force", "update_vof", "density1",
    "mass_flux", "predict", "potential", "correct", "pressure",
]


def run_stage(case: solver.Case, stage: str, k: int, dt: float):
    dt_old, dt_new = solver.substep_dts(dt, k)
    if stage == "density0":
        case.compute_density(0)
    elif stage == "viscosity":
        case.compute_viscosity()
    elif stage == "curvature_tensor":
        case.compute_curvature_tensor()
    elif stage == "force":
        case.compute_force()
    elif stage == "update_vof":
        case.update_vof(k, dt)
    elif stage == "density1":
        case.compute_density(1)
    elif stage == "mass_flux":
        case.compute_mass_flux()
    elif stage == "predict":
        case.predict_field(k, dt)
    elif stage == "potential":
        case.compute_potential(dt_old, dt_new)
    elif stage == "correct":
        case.correct_velocity(dt_old, dt_new)
    elif stage == "pressure":
        case.update_pressure()
    else:
        raise KeyError(stage)
