"""The benchmark and parity workloads.

Physics defaults follow the reference's run script (exec/main.sh:18-30:
coef_dt_adv/dif/int = 0.2/0.5/0.95, Re=35, We=10, Fr=1), with the density
and viscosity ratios 1e-3 / 1e-2 of the workloads below.  Domains are 1 x 2
(x wall-bounded and along gravity, y periodic), so cells are square with
h = 1/Nx.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import ic

DT_COEFS = (0.20, 0.50, 0.95)  # exec/main.sh:21-23


@dataclass
class Workload:
    name: str
    nx: int
    ny: int
    lx: float = 1.0
    ly: float = 2.0
    Re: float = 35.0
    We: float = 10.0
    Fr: float = 1.0
    denr: float = 1.0e-3
    visr: float = 1.0e-2
    steps: int = 1
    bubbles: list = field(default_factory=list)
    seed: int | None = None

    def vof(self) -> np.ndarray:
        if len(self.bubbles) == 1:
            cx, cy, r = self.bubbles[0]
            return ic.bubble_vof(self.nx, self.ny, self.lx, self.ly, cx, cy, r)
        return ic.swarm_vof(self.nx, self.ny, self.lx, self.ly, self.bubbles)

    @property
    def cells(self) -> int:
        return self.nx * self.ny


def rising_bubble(nx: int = 128, ny: int = 256, steps: int = 2000) -> Workload:
    """One bubble of diameter 0.5 centred at (0.5, 0.5) in quiescent liquid
    (128 x 256, 2000 steps in its full form)."""
    return Workload("rising_bubble_%dx%d" % (nx, ny), nx, ny, steps=steps,
                    bubbles=[(0.5, 0.5, 0.25)])


def bubble_swarm(nx: int = 2048, ny: int = 4096, nbubbles: int = 64,
                 steps: int = 200, seed: int = 1, Re: float = 35.0,
                 We: float = 10.0) -> Workload:
    """Non-overlapping bubbles, diameters uniform in [24, 48] cells,
    dart-throwing placement with a 4-cell gap (2048 x 4096 with 64 bubbles
    and 200 steps in its full form)."""
    b = ic.poisson_disk_bubbles(nbubbles, nx, ny, 1.0, 2.0, 24.0, 48.0, 4.0, seed)
    return Workload("bubble_swarm_%dx%d_n%d" % (nx, ny, nbubbles), nx, ny,
                    Re=Re, We=We, steps=steps, bubbles=b, seed=seed)


def static_drop(n: int = 256, steps: int = 5000, radius_cells: float = 32.0) -> Workload:
    """A gas drop of radius 32 cells at the centre of a unit square (n x n),
    no gravity (Fr = 1e30 makes the buoyancy term -1/Fr^2 vanish), at rest:
    any velocity is a spurious current of the surface-tension model."""
    return Workload("static_drop_%dx%d" % (n, n), n, n, lx=1.0, ly=1.0, Fr=1e30, steps=steps,
                    bubbles=[(0.5, 0.5, radius_cells / n)])


def ensemble(n: int = 256, seed: int = 3) -> list:
    """Independent rising bubbles: grids drawn from {64x128, 128x256,
    256x512} on a 1 x 2 domain, density and viscosity ratios log-uniform in
    [1e-3, 1e-1], We uniform in [1, 50], Re uniform in [20, 200], one
    centred bubble of diameter 1/4 of the x extent."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        nx = int(rng.choice([64, 128, 256]))
        w = Workload("ensemble_case_%d_%dx%d" % (k, nx, 2 * nx), nx, 2 * nx, lx=1.0, ly=2.0,
                     Re=float(rng.uniform(20.0, 200.0)), We=float(rng.uniform(1.0, 50.0)), Fr=1.0,
                     denr=float(10.0 ** rng.uniform(-3.0, -1.0)), visr=float(10.0 ** rng.uniform(-3.0, -1.0)),
                     steps=1000, bubbles=[(0.5, 1.0, 0.125)])
        out.append(w)
    return out


def large_swarm() -> Workload:
    """4096 x 8192, 1024 bubbles, Re=100, We=5, 500 steps."""
    return bubble_swarm(4096, 8192, 1024, 500, seed=4, Re=100.0, We=5.0)
