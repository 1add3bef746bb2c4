"""Initial conditions: the reference's IC recipe and the bubble swarms.

The single-bubble field restates initial_condition/main.py:27-43:
    eta = r - |x - c|,  vof = 0.5 (1 + tanh(0.5 Nx eta))
evaluated at the cell centres of x including the two wall points
(main.py:15-24) and the y cell centres, giving an NPY-shaped array of
shape (Ny, Nx+2).  Velocities, pressure and potential start at zero
(main.py:46-61).

Bubble swarms place non-overlapping circles by dart throwing (Bridson-style
rejection sampling, seeded); centres keep a gap between circles, stay clear
of the walls, and wrap periodically in y.  Several bubbles combine by a
per-cell max of their profiles.
"""
from __future__ import annotations

import numpy as np


def grid(nx: int, ny: int, lx: float, ly: float):
    """xf (Nx+1), xc (Nx+2, incl. walls), yc (Ny) as main.py:15-24, 38."""
    xf = np.linspace(0.0, lx, nx + 1, endpoint=True)
    xc = np.concatenate([[0.0], 0.5 * xf[:-1] + 0.5 * xf[1:], [lx]])
    dy = ly / ny
    yc = np.linspace(0.5 * dy, ly - 0.5 * dy, ny)
    return xf, xc, yc


def bubble_vof(nx: int, ny: int, lx: float, ly: float, cx: float, cy: float,
               r: float) -> np.ndarray:
    """One bubble, non-periodic distance, exactly as main.py:27-43."""
    _, xc, yc = grid(nx, ny, lx, ly)
    X, Y = np.meshgrid(xc, yc)
    eta = r - np.sqrt(np.power(X - cx, 2.0) + np.power(Y - cy, 2.0))
    return 0.5 * (1.0 + np.tanh(0.5 * nx * eta))


def swarm_vof(nx: int, ny: int, lx: float, ly: float, bubbles) -> np.ndarray:
    """Per-cell max of the tanh profiles of many bubbles (periodic in y).

    Each profile is evaluated in a window of radius r + 48 cells, beyond
    which 0.5 (1 + tanh(.)) rounds to exactly 0."""
    _, xc, yc = grid(nx, ny, lx, ly)
    h = lx / nx
    vof = np.zeros((ny, nx + 2))
    for (cx, cy, r) in bubbles:
        w = r + 48 * h
        ilo = max(0, int(np.floor((cx - w) / h)))
        ihi = min(nx + 2, int(np.ceil((cx + w) / h)) + 2)
        jlo = int(np.floor((cy - w) / (ly / ny))) - 1
        jhi = int(np.ceil((cy + w) / (ly / ny))) + 1
        js = np.arange(jlo, jhi) % ny
        js = np.unique(js)
        dyv = np.abs(yc[js] - cy)
        dyv = np.minimum(dyv, ly - dyv)
        dxv = xc[ilo:ihi] - cx
        eta = r - np.sqrt(dxv[None, :] ** 2 + dyv[:, None] ** 2)
        prof = 0.5 * (1.0 + np.tanh(0.5 * nx * eta))
        sub = vof[js, ilo:ihi]
        vof[js, ilo:ihi] = np.maximum(sub, prof)
    return vof


def poisson_disk_bubbles(n: int, nx: int, ny: int, lx: float, ly: float,
                         dmin_cells: float, dmax_cells: float, gap_cells: float,
                         seed: int, max_tries: int = 200000):
    """n circles with diameters uniform in [dmin, dmax] cells, centre spacing
    >= r_i + r_j + gap, >= gap cells from the walls, periodic in y."""
    rng = np.random.default_rng(seed)
    h = lx / nx
    out = []
    tries = 0
    while len(out) < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError(f"placed only {len(out)} of {n} bubbles")
        r = 0.5 * rng.uniform(dmin_cells, dmax_cells) * h
        cx = rng.uniform(r + gap_cells * h, lx - r - gap_cells * h)
        cy = rng.uniform(0.0, ly)
        ok = True
        for (bx, by, br) in out:
            dy = abs(cy - by)
            dy = min(dy, ly - dy)
            if (cx - bx) ** 2 + dy ** 2 < (r + br + gap_cells * h) ** 2:
                ok = False
                break
        if ok:
            out.append((cx, cy, r))
    return out


def write_ic_dir(dest, nx: int, ny: int, lx: float, ly: float, vof: np.ndarray,
                 ux: np.ndarray | None = None, uy: np.ndarray | None = None,
                 p: np.ndarray | None = None, psi: np.ndarray | None = None,
                 step: int = 0, time: float = 0.0) -> None:
    """Write an initial-condition directory in the layout the reference's
    generator writes (initial_condition/main.py:6-61): step.npy ('<u8', 0-d),
    time.npy, glsizes.npy ('<u8'), lengths.npy, xf.npy, xc.npy, and the
    fields ux (Ny, Nx+1), uy / p / psi / vof (Ny, Nx+2); absent velocity,
    pressure and potential fields are zero."""
    import os
    os.makedirs(dest, exist_ok=True)
    xf, xc, _ = grid(nx, ny, lx, ly)
    np.save(f"{dest}/step.npy", np.array(step, dtype=np.uint64))
    np.save(f"{dest}/time.npy", np.array(time, dtype=np.float64))
    np.save(f"{dest}/glsizes.npy", np.array([nx, ny], dtype=np.uint64))
    np.save(f"{dest}/lengths.npy", np.array([lx, ly], dtype=np.float64))
    np.save(f"{dest}/xf.npy", np.asarray(xf, dtype=np.float64))
    np.save(f"{dest}/xc.npy", np.asarray(xc, dtype=np.float64))
    s0, s1 = (ny, nx + 1), (ny, nx + 2)
    for name, arr, shape in (("ux", ux, s0), ("uy", uy, s1), ("p", p, s1), ("psi", psi, s1),
                             ("vof", vof, s1)):
        a = np.zeros(shape) if arr is None else np.ascontiguousarray(arr, dtype=np.float64)
        if a.shape != shape:
            raise ValueError(f"{name}: shape {a.shape} != {shape}")
        np.save(f"{dest}/{name}.npy", a)


def main(argv=None) -> int:
    """python -m simplebubblyflowsolver_amd.ic DEST [--workload rising|swarm|large]
    [--nx N --ny N]: write the IC directory of a named workload."""
    import argparse
    from . import configs
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("dest")
    ap.add_argument("--workload", default="rising", choices=["rising", "swarm", "large"])
    ap.add_argument("--nx", type=int)
    ap.add_argument("--ny", type=int)
    a = ap.parse_args(argv)
    if a.workload == "rising":
        w = configs.rising_bubble(a.nx or 128, a.ny or 256)
    elif a.workload == "swarm":
        w = configs.bubble_swarm(a.nx or 2048, a.ny or 4096)
    else:
        w = configs.large_swarm()
    write_ic_dir(a.dest, w.nx, w.ny, w.lx, w.ly, w.vof())
    print(f"{a.dest}: {w.name}, Re={w.Re} We={w.We} Fr={w.Fr} denr={w.denr} visr={w.visr}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
