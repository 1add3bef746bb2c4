"""Closed-form pins of the oracle's momentum sources, stresses, THINC face
quadrature and whole-step composition.

Each check builds a state whose discrete answer the mathematics fixes and
compares the oracle against that answer, computed here in numpy from the
geometry of the staggered grid and the definitions of the method, not from
the oracle's own expressions:

* advection (reference src/fluid/predict/ux.c:29-93, uy.c:29-90): the
  central flux form on uniform cells turns a linear or sinusoidal velocity
  into exact trigonometric / polynomial differences;
* variable-viscosity diffusion and the viscosity placement of the shear
  stress (predict/uy.c:92-118, txy.c:13-50, viscosity.c:13-108);
* the THINC face average (src/interface/update/fluxx.c:53-68, fluxy.c,
  common.c:4-11) against the exact transverse integral of the tanh
  indicator;
* a stratified two-phase column at rest (predict/ux.c:146-192,
  correct_velocity/ux.c:18-66, update_pressure.c:7-53);
* a decaying shear wave, whose velocity after N steps is the third-order
  Runge-Kutta polynomial of the discrete diffusion operator (src/integrate.c,
  runge_kutta.c, predict/uy.c).

Each `check_*` takes the oracle module (or an oracle built from a modified
source, tests/test_oracle_mutations.py) and raises AssertionError when the
oracle departs from the closed form.
"""
import math

import numpy as np
import pytest

from simplebubblyflowsolver_amd import abi

HUGE = 1e300  # We and Fr that switch surface tension and gravity off


def _case(om, nx, ny, lx=1.0, ly=None, **phys):
    ph = dict(Re=35.0, Fr=HUGE, denr=1.0, visr=1.0, We=HUGE)
    ph.update(phys)
    return om.new_case(nx, ny, lx, ly if ly is not None else lx * ny / nx, **ph)


def _grid(c):
    """xf (faces, nx+1), xc (ghost, centres, ghost: nx+2), hy, h."""
    xf, xc = c.metric("xf"), c.metric("xc")
    return xf, xc, c.domain.hy, xf[1] - xf[0]


def _vof_columns(c, v_cells):
    """vof from per-column interior values (uniform in y); wall ghosts and
    the periodic halo by the boundary routine."""
    full = np.zeros((c.ny, c.nx + 2))
    full[:, 1:-1] = v_cells[None, :]
    c.set_interior("vof", full)
    c.update_boundaries("vof")


def _sources(c, dt=1e-3):
    """Run the stages that lead to the momentum sources of Runge-Kutta
    stage 0 (src/integrate.c:18-40) and return the alpha sources."""
    c.compute_density(0)
    c.compute_viscosity()
    c.compute_curvature_tensor()
    c.compute_force()
    c.update_vof(0, dt)
    c.compute_density(1)
    c.compute_mass_flux()
    c.predict_field(0, dt)
    return c.get("srcux_a"), c.get("srcuy_a")


def _ux_rows(c, f_of_j):
    """An ux buffer (rows j = 0..ny+1, faces i = 1..nx+1) filled by rows."""
    ux = np.zeros(c.shape("ux"))
    for j in range(c.ny + 2):
        ux[j] = f_of_j(j)
    return ux


# srcux_a holds faces i = 2..nx (column i - 2) of rows j = 1..ny (row j - 1);
# srcuy_a holds cells i = 1..nx (column i - 1) of rows j = 1..ny.


# ---------------------------------------------------------------------------
# advection
# ---------------------------------------------------------------------------
def check_ux_advection_in_x(om):
    """Single phase, ux = a x, uy = 0.  The face momentum flux of the
    central scheme is (a xc)^2 at each cell centre, so the x advection of
    ux at face i is -(a^2 / h)(xc_i^2 - xc_{i-1}^2) = -2 a^2 xf_i exactly;
    stress divergence vanishes for a linear field and every other source is
    off."""
    nx, ny, a = 32, 32, 0.7
    c = _case(om, nx, ny)
    try:
        xf, xc, hy, h = _grid(c)
        _vof_columns(c, np.zeros(nx))
        c.set("ux", _ux_rows(c, lambda j: a * xf))
        c.set("uy", np.zeros(c.shape("uy")))
        c.set("p", np.zeros(c.shape("p")))
        sx, sy = _sources(c)
        i = np.arange(2, nx)                      # faces whose stencil avoids the right wall
        want = -2 * a * a * xf[i - 1]
        got = sx[:, i - 2]
        assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max(), np.abs(got - want).max()
        assert np.abs(sy).max() <= 1e-15
    finally:
        c.close()


def check_ux_advection_and_diffusion_in_y(om):
    """ux = A sin(k yc) on every face, uy = V on every interior column.
    Central differences of a sinusoid: y advection of ux is
    -V A cos(k yc) sin(k hy) / hy and its y diffusion (unit viscosity)
    (1/Re) A sin(k yc) 2 (cos(k hy) - 1) / hy^2; x terms cancel."""
    nx, ny, A, V, Re = 16, 32, 0.4, 0.3, 20.0
    c = _case(om, nx, ny, Re=Re)
    try:
        xf, xc, hy, h = _grid(c)
        k = 2 * math.pi * 2 / (ny * hy)
        yc = (np.arange(ny + 2) - 0.5) * hy          # centre of row j = 0..ny+1
        _vof_columns(c, np.zeros(nx))
        c.set("ux", _ux_rows(c, lambda j: np.full(nx + 1, A * math.sin(k * yc[j]))))
        uy = np.full(c.shape("uy"), V)
        uy[:, 0] = uy[:, -1] = 0.0
        c.set("uy", uy)
        c.set("p", np.zeros(c.shape("p")))
        sx, _ = _sources(c)
        y = yc[1:ny + 1][:, None]
        want = (-V * A * np.cos(k * y) * math.sin(k * hy) / hy
                + (1 / Re) * A * np.sin(k * y) * 2 * (math.cos(k * hy) - 1) / hy ** 2)
        want = np.broadcast_to(want, sx.shape)
        assert np.abs(sx - want).max() <= 1e-12 * np.abs(want).max(), np.abs(sx - want).max()
    finally:
        c.close()


def check_uy_advection_and_diffusion_in_x(om):
    """uy = B sin(pi x) at the cell centres (zero at both walls, matching
    the wall values), ux = U on every interior face.  x advection of uy is
    -U B cos(pi xc) sin(pi h) / h and its x diffusion
    (1/Re) B sin(pi xc) 2 (cos(pi h) - 1) / h^2 for cells away from the
    walls."""
    nx, ny, B, U, Re = 32, 32, 0.5, -0.25, 15.0
    c = _case(om, nx, ny, Re=Re)
    try:
        xf, xc, hy, h = _grid(c)
        _vof_columns(c, np.zeros(nx))
        ux = np.full(c.shape("ux"), U)
        ux[:, 0] = ux[:, -1] = 0.0
        c.set("ux", ux)
        row = B * np.sin(math.pi * xc)
        row[0] = row[-1] = 0.0
        c.set("uy", np.tile(row, (ny + 2, 1)))
        c.set("p", np.zeros(c.shape("p")))
        _, sy = _sources(c)
        i = np.arange(2, nx)
        x = xc[i]
        want = (-U * B * np.cos(math.pi * x) * math.sin(math.pi * h) / h
                + (1 / Re) * B * np.sin(math.pi * x) * 2 * (math.cos(math.pi * h) - 1) / h ** 2)
        got = sy[:, i - 1]
        assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max(), np.abs(got - want).max()
    finally:
        c.close()


def check_uy_advection_and_diffusion_in_y(om):
    """uy = V + A sin(k yf) on the interior columns, ux = 0.  With the face
    means um, up of the two neighbouring cell centres, the y advection is
    -(up^2 - um^2) / hy, i.e.
    -A cos(k yf) sin(k hy) (2 V + A sin(k yf)(1 + cos(k hy))) / hy, and the
    normal-stress diffusion (2/Re) A sin(k yf) 2 (cos(k hy) - 1) / hy^2."""
    nx, ny, A, V, Re = 16, 32, 0.3, 0.2, 25.0
    c = _case(om, nx, ny, Re=Re)
    try:
        xf, xc, hy, h = _grid(c)
        k = 2 * math.pi * 3 / (ny * hy)
        yf = (np.arange(ny + 2) - 1.0) * hy          # face of row j = 0..ny+1
        _vof_columns(c, np.zeros(nx))
        c.set("ux", np.zeros(c.shape("ux")))
        uy = np.zeros(c.shape("uy"))
        for j in range(ny + 2):
            uy[j, 1:-1] = V + A * math.sin(k * yf[j])
        c.set("uy", uy)
        c.set("p", np.zeros(c.shape("p")))
        _, sy = _sources(c)
        y = yf[1:ny + 1][:, None]
        want = (-A * np.cos(k * y) * math.sin(k * hy) * (2 * V + A * np.sin(k * y) * (1 + math.cos(k * hy))) / hy
                + (2 / Re) * A * np.sin(k * y) * 2 * (math.cos(k * hy) - 1) / hy ** 2)
        want = np.broadcast_to(want, (ny, nx - 2))
        got = sy[:, 1:nx - 1]
        assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max(), np.abs(got - want).max()
    finally:
        c.close()


# ---------------------------------------------------------------------------
# variable viscosity
# ---------------------------------------------------------------------------
def check_variable_viscosity_diffusion(om):
    """vof = c0 + c1 x (inside [0, 1], so the mixture viscosity
    mu = 1 + (visr - 1) vof is linear and never clamped) and uy = b x^2,
    ux = 0.  The flux-form x diffusion of uy is
    (1/Re) d/dx(mu duy/dx) = (2 b / Re)(mu(xc) + xc mu'), which the
    discrete operator reproduces exactly for these polynomials."""
    nx, ny, b, Re, visr = 32, 32, 0.9, 12.0, 0.3
    c0, c1 = 0.2, 0.6
    c = _case(om, nx, ny, Re=Re, visr=visr)
    try:
        xf, xc, hy, h = _grid(c)
        _vof_columns(c, c0 + c1 * xc[1:-1])
        c.set("ux", np.zeros(c.shape("ux")))
        row = b * xc ** 2
        row[0] = row[-1] = 0.0
        c.set("uy", np.tile(row, (ny + 2, 1)))
        c.set("p", np.zeros(c.shape("p")))
        _, sy = _sources(c)
        i = np.arange(2, nx)
        x = xc[i]
        mu = 1 + (visr - 1) * (c0 + c1 * x)
        dmu = (visr - 1) * c1
        want = (2 * b / Re) * (mu + x * dmu)
        got = sy[:, i - 1]
        assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max(), np.abs(got - want).max()
    finally:
        c.close()


def check_shear_stress_uses_x_face_viscosity(om):
    """Where the ux-gradient part of the shear stress lives.  With
    ux = A sin(k yc) and uy = 0, txy at the corner (xf_i, yf_j) is
    mu_f(i) (ux_j - ux_{j-1}) / hy, where mu_f(i) is the viscosity of the
    x face i: the clamped mixture rule of the face-averaged vof
    (viscosity.c:25-60), averaged over the two faces above and below the
    corner.  A vof alternating between -0.4 and 0.6 across columns is
    clamped in every other cell, so this differs from the mean of the
    clamped cell viscosities that the uy-gradient part carries.  The y
    diffusion of ux follows as (1/Re) mu_f A sin(k yc) 2(cos(k hy)-1)/hy^2."""
    nx, ny, A, Re, visr = 16, 32, 0.5, 10.0, 0.05
    c = _case(om, nx, ny, Re=Re, visr=visr)
    try:
        xf, xc, hy, h = _grid(c)
        k = 2 * math.pi / (ny * hy)
        yc = (np.arange(ny + 2) - 0.5) * hy
        v = np.where(np.arange(nx) % 2 == 0, -0.4, 0.6)
        _vof_columns(c, v)
        c.set("ux", _ux_rows(c, lambda j: np.full(nx + 1, A * math.sin(k * yc[j]))))
        c.set("uy", np.zeros(c.shape("uy")))
        c.set("p", np.zeros(c.shape("p")))
        sx, _ = _sources(c)
        lo, hi = min(1.0, visr), max(1.0, visr)
        mu_face = np.clip(1 + (visr - 1) * 0.5 * (v[:-1] + v[1:]), lo, hi)      # faces 2..nx
        mu_cell = np.clip(1 + (visr - 1) * v, lo, hi)
        mu_other = 0.5 * (mu_cell[:-1] + mu_cell[1:])
        assert np.abs(mu_face - mu_other).min() > 0.1 * visr                 # the two placements differ
        txy = c.get("txy")                                                   # columns i = 1..nx+1
        j = np.arange(1, ny + 1)[:, None]
        dux = A * (np.sin(k * yc[j]) - np.sin(k * yc[j - 1])) / hy
        want_t = mu_face[None, :] * dux
        got_t = txy[1:ny + 1, 1:nx]                                          # corners i = 2..nx
        assert np.abs(got_t - want_t).max() <= 1e-13 * np.abs(want_t).max(), np.abs(got_t - want_t).max()
        y = yc[1:ny + 1][:, None]
        want = (1 / Re) * mu_face[None, :] * A * np.sin(k * y) * 2 * (math.cos(k * hy) - 1) / hy ** 2
        assert np.abs(sx - want).max() <= 1e-12 * np.abs(want).max(), np.abs(sx - want).max()
    finally:
        c.close()


# ---------------------------------------------------------------------------
# THINC face quadrature
# ---------------------------------------------------------------------------
def thinc_face_average(a, b):
    """Exact mean over s in [-1/2, 1/2] of 1/2 (1 + tanh(a + b s)):
    1/2 + (ln cosh(a + b/2) - ln cosh(a - b/2)) / (2 b)."""
    def lncosh(z):
        z = np.abs(z)
        return z + np.log1p(np.exp(-2 * z)) - math.log(2.0)
    return 0.5 + (lncosh(a + b / 2) - lncosh(a - b / 2)) / (2 * b)


# Two-point Gauss-Legendre is exact for cubics; on these faces its error
# f''''/4320 stays below 2.2e-4, while a rule through the face ends (+-1/2)
# errs by up to 2e-2 and the midpoint rule by up to 1e-2.
THINC_TOL = 1e-3


def check_thinc_oblique_face_average(om):
    """A vof linear in x and y on square cells, v = 0.1 + 0.3 x + 0.2 y:
    the cell normal is exactly (0.6, 0.4) / |(0.6, 0.4)| and the intercept
    d = atanh(2 v - 1) (curvature_tensor.c:53-95).  The flux through a face
    is u times the face average of the upwind cell's indicator
    1/2 (1 + tanh(n . x + d)), with the transverse coordinate integrated
    over the face; compare with the exact integral, for both velocity
    signs and both directions."""
    nx, ny = 32, 64                        # 1 x 2 domain: h = hy = 1/32
    n = np.array([0.6, 0.4]) / math.hypot(0.6, 0.4)
    for axis in (0, 1):
        for U in (0.35, -0.35):
            c = _case(om, nx, ny, denr=0.1)
            try:
                xf, xc, hy, h = _grid(c)
                yc = (np.arange(1, ny + 1) - 0.5) * hy
                v = 0.1 + 0.3 * xc[None, 1:-1] + 0.2 * yc[:, None]
                full = np.zeros((ny, nx + 2))
                full[:, 1:-1] = v
                c.set_interior("vof", full)
                c.update_boundaries("vof")
                c.compute_curvature_tensor()
                if axis == 0:
                    ux = np.full(c.shape("ux"), U)
                    ux[:, 0] = ux[:, -1] = 0.0
                    c.set("ux", ux)
                    c.set("uy", np.zeros(c.shape("uy")))
                else:
                    c.set("ux", np.zeros(c.shape("ux")))
                    uy = np.full(c.shape("uy"), U)
                    uy[:, 0] = uy[:, -1] = 0.0
                    c.set("uy", uy)
                c.update_vof(0, 1e-9)
                rows = np.arange(8, ny - 8)            # away from the periodic seam of the ramp
                cols = np.arange(5, nx - 4)            # away from the walls' ghost vof
                side = 0.5 if U > 0 else -0.5          # downstream face of the upwind cell
                if axis == 0:
                    F = c.get("fluxx")                 # face i at column i - 1, row j at row j
                    up = cols - 1 if U > 0 else cols   # upwind cell index
                    vv = v[rows[:, None] - 1, up[None, :] - 1]
                    got = F[rows[:, None], cols[None, :] - 1] / U
                    a, b = n[0] * side + np.arctanh(2 * vv - 1), n[1]
                else:
                    F = c.get("fluxy")                 # face j at row j, cell i at column i
                    up = rows - 1 if U > 0 else rows
                    vv = v[up[:, None] - 1, cols[None, :] - 1]
                    got = F[rows[:, None], cols[None, :]] / U
                    a, b = n[1] * side + np.arctanh(2 * vv - 1), n[0]
                want = thinc_face_average(a, b)
                err = np.abs(got - want).max()
                assert err <= THINC_TOL, (axis, U, err)
            finally:
                c.close()


# ---------------------------------------------------------------------------
# composition: a column at rest, a decaying shear wave
# ---------------------------------------------------------------------------
def check_stratified_column_at_rest(om, denr=0.1, steps=80):
    """Gas over liquid, stratified along x (gravity is along -x), no
    variation in y.  The discrete divergence constraint with zero wall
    velocity leaves ux = 0 as the only solenoidal field, so every step must
    keep the fluid at rest (to round-off), and the pressure must converge
    to hydrostatic balance dp/dx = -rho_face / Fr^2 on every face whose
    stencil does not reach the gas-filled top wall (whose zero ghost vof
    gives a wall curvature and a surface force in the reference's scheme)."""
    nx, ny, Fr = 32, 32, 1.0
    c = _case(om, nx, ny, Re=35.0, Fr=Fr, denr=denr, visr=1e-2, We=10.0)
    try:
        xf, xc, hy, h = _grid(c)
        v = 0.5 * (1 + np.tanh(8 * (xc[1:-1] - 0.5)))
        full = np.zeros((ny, nx + 2))
        full[:, 1:-1] = v
        c.load_initial_condition(full)
        rho = 1 + (denr - 1) * v
        rho_f = 0.5 * rho[:-1] + 0.5 * rho[1:]               # interior faces 2..nx
        res = []
        for _ in range(steps):
            c.integrate()
            umax = max(np.abs(c.get("ux")).max(), np.abs(c.get("uy")).max())
            assert umax <= 1e-15, umax
            p = abi.interior("p", c.get("p"), nx, ny)
            r = (p[:, 1:] - p[:, :-1]) / h + rho_f / Fr ** 2
            res.append(np.abs(r[:, :-3]).max())
        assert res[-1] <= 1e-4 * (1 / Fr ** 2), res[-1]
        assert res[-1] < 0.05 * res[0]
    finally:
        c.close()


def diffusion_operator(c, Re):
    """(1/Re) d2/dx2 of the cell-centred uy with the wall values (zero) held
    at the wall points: second-order finite volumes on cells of width h whose
    centre-to-wall distance is h/2."""
    nx = c.nx
    xf, xc, hy, h = _grid(c)
    d = np.diff(xc)                                  # centre-to-centre (wall: h/2), nx + 1 faces
    L = np.zeros((nx, nx))
    for i in range(nx):
        w = xf[i + 1] - xf[i]
        L[i, i] = -(1 / d[i] + 1 / d[i + 1]) / w
        if i > 0:
            L[i, i - 1] = 1 / d[i] / w
        if i < nx - 1:
            L[i, i + 1] = 1 / d[i + 1] / w
    return L / Re


def check_decaying_shear_wave(om, steps=12, rtol=1e-13):
    """Single phase, uy = uy(x) only, ux = 0: advection and the pressure
    vanish identically, so the step integrates the linear system
    duy/dt = L uy with L the discrete viscous operator.  Any three-stage
    third-order Runge-Kutta scheme advances a linear system by
    P(dt L) = I + Z + Z^2/2 + Z^3/6 per step, so after N steps
    uy = prod P(dt_n L) uy_0 exactly (the dt of each step is what
    integrate returns)."""
    nx, ny, Re = 32, 32, 20.0
    c = _case(om, nx, ny, Re=Re)
    try:
        xf, xc, hy, h = _grid(c)
        u0 = 0.1 * np.sin(math.pi * xc[1:-1]) + 0.05 * np.sin(3 * math.pi * xc[1:-1]) \
            + 0.02 * np.cos(5 * math.pi * xc[1:-1])
        row = np.zeros(nx + 2)
        row[1:-1] = u0
        _vof_columns(c, np.zeros(nx))
        c.set("ux", np.zeros(c.shape("ux")))
        c.set("uy", np.tile(row, (ny + 2, 1)))
        c.set("p", np.zeros(c.shape("p")))
        c.set("psi0", np.zeros(c.shape("psi0")))
        L = diffusion_operator(c, Re)
        I = np.eye(nx)
        u = u0.copy()
        for _ in range(steps):
            dt = c.integrate()
            Z = dt * L
            u = (I + Z + Z @ Z / 2 + Z @ Z @ Z / 6) @ u
        got = abi.interior("uy", c.get("uy"), nx, ny)
        assert np.abs(got - u[None, :]).max() <= rtol * np.abs(u0).max(), np.abs(got - u[None, :]).max()
        assert np.abs(u - u0).max() > 1e-3 * np.abs(u0).max()   # the wave did decay measurably
        assert np.abs(c.get("ux")).max() == 0.0
        p = abi.interior("p", c.get("p"), nx, ny)
        assert np.abs(p - p.mean()).max() <= 1e-15
    finally:
        c.close()


CHECKS = {
    "ux_advection_in_x": check_ux_advection_in_x,
    "ux_advection_and_diffusion_in_y": check_ux_advection_and_diffusion_in_y,
    "uy_advection_and_diffusion_in_x": check_uy_advection_and_diffusion_in_x,
    "uy_advection_and_diffusion_in_y": check_uy_advection_and_diffusion_in_y,
    "variable_viscosity_diffusion": check_variable_viscosity_diffusion,
    "shear_stress_uses_x_face_viscosity": check_shear_stress_uses_x_face_viscosity,
    "thinc_oblique_face_average": check_thinc_oblique_face_average,
    "stratified_column_at_rest": check_stratified_column_at_rest,
    "decaying_shear_wave": check_decaying_shear_wave,
}


@pytest.mark.parametrize("name", list(CHECKS))
def test_closed_form(oracle_mod, name):
    CHECKS[name](oracle_mod)


def test_thinc_tolerance_separates_quadrature_rules():
    """The tolerance above passes the two-point Gauss rule and fails rules
    with other nodes on the faces the THINC check visits (scipy quad is the
    yardstick for the closed-form average)."""
    from scipy.integrate import quad
    n = np.array([0.6, 0.4]) / math.hypot(0.6, 0.4)
    g = 1 / (2 * math.sqrt(3))
    worst = {"gauss": 0.0, "ends": 0.0, "mid": 0.0}
    for v in np.linspace(0.05, 0.95, 61):
        d = math.atanh(2 * v - 1)
        for side in (0.5, -0.5):
            for nn, nt in ((n[0], n[1]), (n[1], n[0])):
                a = nn * side + d
                exact = float(thinc_face_average(np.array(a), nt))
                q, _ = quad(lambda s: 0.5 * (1 + math.tanh(a + nt * s)), -0.5, 0.5, epsabs=1e-14)
                assert abs(q - exact) <= 1e-13
                f = lambda s: 0.5 * (1 + math.tanh(a + nt * s))  # noqa: E731
                worst["gauss"] = max(worst["gauss"], abs(0.5 * (f(-g) + f(g)) - exact))
                worst["ends"] = max(worst["ends"], abs(0.5 * (f(-0.5) + f(0.5)) - exact))
                worst["mid"] = max(worst["mid"], abs(f(0.0) - exact))
    assert worst["gauss"] < THINC_TOL / 4
    assert worst["ends"] > 10 * THINC_TOL and worst["mid"] > 5 * THINC_TOL
