"""Host-side mirror of the reference's solver interface.

`Backend` binds a shared library that exports the reference's entry points
(integrate, decide_dt, fluid_*, interface_*; reference include/integrate.h,
decide_dt.h, fluid_solver.h, interface_solver.h).  `Case` owns one
domain_t / fluid_t / interface_t triple and exposes the stages by the
reference's names, so a test reads like a call sequence of src/integrate.c.

The product backend is libsbfs.so (HIP, arrays in device memory).  The same
classes also drive the CPU oracle (oracle/liboracle.so, arrays in host
memory) for parity checks.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

from . import abi

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libsbfs.so"

_P = C.POINTER
_STAGES = {
    # name: argtypes (after the leading domain pointer)
    "integrate": [_P(abi.FluidT), _P(abi.InterfaceT), _P(C.c_double)],
    "decide_dt": [_P(abi.FluidT), _P(abi.InterfaceT), _P(C.c_double)],
    "fluid_compute_density": [_P(abi.FluidT), _P(abi.InterfaceT), C.c_size_t],
    "fluid_compute_viscosity": [_P(abi.FluidT), _P(abi.InterfaceT)],
    "interface_compute_curvature_tensor": [_P(abi.InterfaceT)],
    "interface_compute_force": [_P(abi.FluidT), _P(abi.InterfaceT)],
    "interface_update_vof": [C.c_size_t, C.c_double, _P(abi.FluidT), _P(abi.InterfaceT)],
    "interface_compute_mass_flux": [_P(abi.FluidT), _P(abi.InterfaceT)],
    "fluid_predict_field": [C.c_size_t, C.c_double, _P(abi.FluidT), _P(abi.InterfaceT)],
    "fluid_compute_potential": [C.c_double, C.c_double, _P(abi.FluidT)],
    "fluid_correct_velocity": [C.c_double, C.c_double, _P(abi.FluidT)],
    "fluid_update_pressure": [_P(abi.FluidT)],
    "fluid_update_boundaries_ux": [_P(abi.ArrayT)],
    "fluid_update_boundaries_uy": [_P(abi.ArrayT)],
    "fluid_update_boundaries_p": [_P(abi.ArrayT)],
    "fluid_update_boundaries_psi": [_P(abi.ArrayT)],
    "interface_update_boundaries_vof": [_P(abi.ArrayT)],
}


class SolverError(RuntimeError):
    pass


def load_library(path: os.PathLike | str | None = None) -> C.CDLL:
    """Load the HIP library.  Raises if it has not been built: there is no
    CPU fallback for the product path."""
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise SolverError(
            f"{p} not found: build the HIP library first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    return C.CDLL(str(p))


class Backend:
    """A library exporting the reference stage API.

    stage_prefix   prefix of the stage symbols ("" for libsbfs.so)
    create_prefix  prefix of the *_create / *_destroy helpers
    device_memory  True if array_t.data holds device pointers
    """

    def __init__(self, lib: C.CDLL, stage_prefix: str, create_prefix: str,
                 device_memory: bool, error_fn: str | None = None):
        self.lib = lib
        self.device_memory = device_memory
        self._fns = {}
        for name, args in _STAGES.items():
            f = getattr(lib, stage_prefix + name)
            f.argtypes = [_P(abi.DomainT)] + args
            f.restype = C.c_int
            self._fns[name] = f
        cp = create_prefix
        self.domain_create = getattr(lib, cp + "domain_create")
        self.domain_create.argtypes = [C.c_size_t, C.c_size_t, C.c_double, C.c_double, _P(abi.DomainT)]
        self.domain_destroy = getattr(lib, cp + "domain_destroy")
        self.domain_destroy.argtypes = [_P(abi.DomainT)]
        self.fluid_create = getattr(lib, cp + "fluid_create")
        self.fluid_create.argtypes = [_P(abi.DomainT), C.c_double, C.c_double, C.c_double, C.c_double, _P(abi.FluidT)]
        self.fluid_destroy = getattr(lib, cp + "fluid_destroy")
        self.fluid_destroy.argtypes = [_P(abi.FluidT)]
        self.interface_create = getattr(lib, cp + "interface_create")
        self.interface_create.argtypes = [_P(abi.DomainT), C.c_double, _P(abi.InterfaceT)]
        self.interface_destroy = getattr(lib, cp + "interface_destroy")
        self.interface_destroy.argtypes = [_P(abi.InterfaceT)]
        self.diagnostics = getattr(lib, cp + "diagnostics")
        self.diagnostics.argtypes = [_P(abi.DomainT), _P(abi.FluidT), _P(abi.InterfaceT), _P(C.c_double)]
        self.set_dt_coefs = getattr(lib, cp + "set_dt_coefs")
        self.set_dt_coefs.argtypes = [C.c_double, C.c_double, C.c_double]
        self._err = None
        if error_fn:
            self._err = getattr(lib, error_fn)
            self._err.restype = C.c_char_p
        if device_memory:
            self.upload = lib.sbfs_array_upload
            self.upload.argtypes = [_P(abi.ArrayT), C.c_void_p, C.c_size_t]
            self.download = lib.sbfs_array_download
            self.download.argtypes = [_P(abi.ArrayT), C.c_void_p, C.c_size_t]
            self.synchronize = lib.sbfs_synchronize

    def error_text(self) -> str:
        return self._err().decode() if self._err else ""

    def call(self, name: str, *args) -> None:
        rc = self._fns[name](*args)
        if rc != 0:
            raise SolverError(f"{name} returned {rc}: {self.error_text()}")

    def check(self, rc: int, what: str) -> None:
        if rc != 0:
            raise SolverError(f"{what} returned {rc}: {self.error_text()}")


def product_backend(path: os.PathLike | str | None = None) -> Backend:
    lib = load_library(path)
    return Backend(lib, "", "sbfs_", True, error_fn="sbfs_last_error")


class CurvatureNet:
    """The learned-curvature network of libsbfs.so (include/sbfs.h)."""

    PLAIN, LEARNED = 0, 1

    def __init__(self, backend: Backend):
        self.be = backend
        lib = backend.lib
        f = C.c_float
        lib.sbfs_mlp_num_params.restype = C.c_int
        self.num_params = int(lib.sbfs_mlp_num_params())
        lib.sbfs_mlp_train.argtypes = [C.c_uint64, C.c_size_t, C.c_int, C.c_int, C.c_double, _P(C.c_double)]
        lib.sbfs_mlp_set_params.argtypes = [_P(f), C.c_size_t]
        lib.sbfs_mlp_get_params.argtypes = [_P(f), C.c_size_t]
        lib.sbfs_mlp_samples.argtypes = [C.c_uint64, C.c_size_t, C.c_size_t, _P(f), _P(f), _P(f), _P(f), _P(f)]
        lib.sbfs_mlp_gradient.argtypes = [_P(f), _P(f), _P(f), C.c_size_t, _P(f), _P(C.c_double)]
        lib.sbfs_mlp_predict.argtypes = [_P(C.c_double), C.c_size_t, _P(C.c_double)]
        lib.sbfs_set_curvature.argtypes = [C.c_int]
        lib.sbfs_mlp_coverage.argtypes = [_P(abi.DomainT), _P(abi.InterfaceT), _P(C.c_double)]
        self.lib = lib

    @staticmethod
    def _p(a, t):
        return a.ctypes.data_as(_P(t))

    def train(self, seed: int, nsamples: int, epochs: int, batch: int, lr: float) -> np.ndarray:
        hist = np.zeros(epochs)
        self.be.check(self.lib.sbfs_mlp_train(seed, nsamples, epochs, batch, lr, self._p(hist, C.c_double)),
                      "sbfs_mlp_train")
        return hist

    def params(self) -> np.ndarray:
        p = np.zeros(self.num_params, np.float32)
        self.be.check(self.lib.sbfs_mlp_get_params(self._p(p, C.c_float), p.size), "sbfs_mlp_get_params")
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        self.be.check(self.lib.sbfs_mlp_set_params(self._p(p, C.c_float), p.size), "sbfs_mlp_set_params")

    def samples(self, seed: int, first: int, n: int):
        """(features [n][32], labels, stencils [n][25], targets, weights)."""
        x = np.zeros((n, 32), np.float32)
        lab = np.zeros(n, np.float32)
        t = np.zeros(n, np.float32)
        w = np.zeros(n, np.float32)
        st = np.zeros((n, 25), np.float32)
        self.be.check(self.lib.sbfs_mlp_samples(seed, first, n, self._p(x, C.c_float), self._p(lab, C.c_float),
                                                self._p(t, C.c_float), self._p(w, C.c_float),
                                                self._p(st, C.c_float)), "sbfs_mlp_samples")
        return x, lab, st, t, w

    def gradient(self, x: np.ndarray, t: np.ndarray, w: np.ndarray | None = None):
        """Gradient of mean(w (y - t)^2) at the current parameters."""
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        wp = None
        if w is not None:
            w = np.ascontiguousarray(w, np.float32)
            wp = self._p(w, C.c_float)
        g = np.zeros(self.num_params, np.float32)
        loss = C.c_double(0.0)
        self.be.check(self.lib.sbfs_mlp_gradient(self._p(x, C.c_float), self._p(t, C.c_float), wp, t.size,
                                                 self._p(g, C.c_float), C.byref(loss)), "sbfs_mlp_gradient")
        return loss.value, g

    def predict(self, stencils: np.ndarray) -> np.ndarray:
        st = np.ascontiguousarray(stencils, np.float64).reshape(-1, 25)
        y = np.zeros(st.shape[0])
        self.be.check(self.lib.sbfs_mlp_predict(self._p(st, C.c_double), st.shape[0], self._p(y, C.c_double)),
                      "sbfs_mlp_predict")
        return y

    def coverage(self, case: "Case") -> float:
        """Fraction of the case's interior cells where the learned value is taken."""
        f = C.c_double(0.0)
        self.be.check(self.lib.sbfs_mlp_coverage(C.byref(case.domain), C.byref(case.interface), C.byref(f)),
                      "sbfs_mlp_coverage")
        return f.value

    def use(self, learned: bool) -> None:
        self.be.check(self.lib.sbfs_set_curvature(self.LEARNED if learned else self.PLAIN), "sbfs_set_curvature")


class Case:
    """One simulation: domain_t + fluid_t + interface_t on a backend."""

    def __init__(self, backend: Backend, nx: int, ny: int, lx: float, ly: float,
                 Re: float, Fr: float, denr: float, visr: float, We: float):
        self.be = backend
        self.nx, self.ny = int(nx), int(ny)
        self.domain = abi.DomainT()
        self.fluid = abi.FluidT()
        self.interface = abi.InterfaceT()
        be = backend
        be.check(be.domain_create(self.nx, self.ny, lx, ly, C.byref(self.domain)), "domain_create")
        be.check(be.fluid_create(C.byref(self.domain), Re, Fr, denr, visr, C.byref(self.fluid)), "fluid_create")
        be.check(be.interface_create(C.byref(self.domain), We, C.byref(self.interface)), "interface_create")
        self._open = True

    @classmethod
    def view(cls, backend: Backend, domain: int, fluid: int, interface: int) -> "Case":
        """A case over structs another object owns (an ensemble member):
        the addresses of its domain_t, fluid_t and interface_t."""
        self = cls.__new__(cls)
        self.be = backend
        self.domain = abi.DomainT.from_address(domain)
        self.fluid = abi.FluidT.from_address(fluid)
        self.interface = abi.InterfaceT.from_address(interface)
        self.nx, self.ny = int(self.domain.mysizes[0]), int(self.domain.mysizes[1])
        self._open = False
        return self

    # -- lifetime ---------------------------------------------------------
    def close(self) -> None:
        if self._open:
            self.be.interface_destroy(C.byref(self.interface))
            self.be.fluid_destroy(C.byref(self.fluid))
            self.be.domain_destroy(C.byref(self.domain))
            self._open = False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- metrics ----------------------------------------------------------
    def metric(self, name: str) -> np.ndarray:
        n = {"xf": self.nx + 1, "xc": self.nx + 2, "hxxf": self.nx + 1,
             "hxxc": self.nx, "jdxf": self.nx + 1, "jdxc": self.nx}[name]
        return np.ctypeslib.as_array(getattr(self.domain, name), shape=(n,)).copy()

    # -- fields -----------------------------------------------------------
    def array(self, name: str) -> abi.ArrayT:
        return abi.get_array(self.fluid, self.interface, name)

    def shape(self, name: str) -> tuple:
        return abi.field_shape(name, self.nx, self.ny)

    def get(self, name: str) -> np.ndarray:
        a = self.array(name)
        out = np.empty(self.shape(name), dtype=np.float64)
        assert out.nbytes == a.datasize, (name, out.nbytes, a.datasize)
        if self.be.device_memory:
            self.be.check(self.be.download(C.byref(a), out.ctypes.data, out.nbytes), f"download {name}")
        else:
            C.memmove(out.ctypes.data, a.data, out.nbytes)
        return out

    def set(self, name: str, values: np.ndarray) -> None:
        a = self.array(name)
        v = np.ascontiguousarray(values, dtype=np.float64)
        if v.shape != self.shape(name):
            raise ValueError(f"{name}: shape {v.shape} != {self.shape(name)}")
        if self.be.device_memory:
            self.be.check(self.be.upload(C.byref(a), v.ctypes.data, v.nbytes), f"upload {name}")
        else:
            C.memmove(a.data, v.ctypes.data, v.nbytes)

    def set_interior(self, name: str, values: np.ndarray) -> None:
        """Set the y-interior rows (all x entries incl. ghosts), as an NPY
        initial-condition file holds them (reference src/array.c:190-254)."""
        full = self.get(name)
        r0, _ = abi.field_origin(name)
        full[r0:r0 + self.ny] = values
        self.set(name, full)

    # -- stages (reference names) -----------------------------------------
    def _d(self):
        return C.byref(self.domain)

    def integrate(self) -> float:
        dt = C.c_double(0.0)
        self.be.call("integrate", self._d(), C.byref(self.fluid), C.byref(self.interface), C.byref(dt))
        return dt.value

    def decide_dt(self) -> float:
        dt = C.c_double(0.0)
        self.be.call("decide_dt", self._d(), C.byref(self.fluid), C.byref(self.interface), C.byref(dt))
        return dt.value

    def compute_density(self, index: int) -> None:
        self.be.call("fluid_compute_density", self._d(), C.byref(self.fluid), C.byref(self.interface), index)

    def compute_viscosity(self) -> None:
        self.be.call("fluid_compute_viscosity", self._d(), C.byref(self.fluid), C.byref(self.interface))

    def compute_curvature_tensor(self) -> None:
        self.be.call("interface_compute_curvature_tensor", self._d(), C.byref(self.interface))

    def compute_force(self) -> None:
        self.be.call("interface_compute_force", self._d(), C.byref(self.fluid), C.byref(self.interface))

    def update_vof(self, rkstep: int, dt: float) -> None:
        self.be.call("interface_update_vof", self._d(), rkstep, dt, C.byref(self.fluid), C.byref(self.interface))

    def compute_mass_flux(self) -> None:
        self.be.call("interface_compute_mass_flux", self._d(), C.byref(self.fluid), C.byref(self.interface))

    def predict_field(self, rkstep: int, dt: float) -> None:
        self.be.call("fluid_predict_field", self._d(), rkstep, dt, C.byref(self.fluid), C.byref(self.interface))

    def compute_potential(self, dt_old: float, dt_new: float) -> None:
        self.be.call("fluid_compute_potential", self._d(), dt_old, dt_new, C.byref(self.fluid))

    def correct_velocity(self, dt_old: float, dt_new: float) -> None:
        self.be.call("fluid_correct_velocity", self._d(), dt_old, dt_new, C.byref(self.fluid))

    def update_pressure(self) -> None:
        self.be.call("fluid_update_pressure", self._d(), C.byref(self.fluid))

    def update_boundaries(self, name: str) -> None:
        fn = {"ux": "fluid_update_boundaries_ux", "uy": "fluid_update_boundaries_uy",
              "p": "fluid_update_boundaries_p", "psi0": "fluid_update_boundaries_psi",
              "psi1": "fluid_update_boundaries_psi",
              "vof": "interface_update_boundaries_vof"}[name]
        self.be.call(fn, self._d(), C.byref(self.array(name)))

    def diagnostics(self) -> dict:
        out = (C.c_double * 8)()
        self.be.check(self.be.diagnostics(self._d(), C.byref(self.fluid), C.byref(self.interface), out),
                      "diagnostics")
        keys = ("volume", "centroid", "rise_velocity", "divmax", "vofmin", "vofmax", "vofmean", "area")
        return dict(zip(keys, list(out)))

    # -- initial condition ------------------------------------------------
    def load_initial_condition(self, vof: np.ndarray, ux: np.ndarray | None = None,
                               uy: np.ndarray | None = None, p: np.ndarray | None = None,
                               psi: np.ndarray | None = None) -> None:
        """Fill the state from NPY-shaped interior arrays and apply the
        boundary conditions, as fluid_init / interface_init do
        (reference src/fluid/init.c:92-100, src/interface/init.c:62-64)."""
        self.set_interior("vof", vof)
        self.update_boundaries("vof")
        for name, v in (("ux", ux), ("uy", uy), ("p", p), ("psi0", psi)):
            if v is not None:
                self.set_interior(name, v)
            self.update_boundaries(name)


RK = [[32. / 60., 0. / 60., 32. / 60. + 0. / 60.],
      [25. / 60., -17. / 60., 25. / 60. + -17. / 60.],
      [45. / 60., -25. / 60., 45. / 60. + -25. / 60.]]


def substep_dts(dt: float, rkstep: int) -> tuple:
    """(dt_old, dt_new) of one Runge-Kutta stage (reference src/integrate.c:54-55)."""
    return RK[(rkstep + 2) % 3][2] * dt, RK[rkstep][2] * dt


def staged_step(case: Case) -> float:
    """One time step as the reference's integrate() composes it
    (src/integrate.c:11-70), issued stage by stage from the host."""
    dt = case.decide_dt()
    for k in range(3):
        case.compute_density(0)
        case.compute_viscosity()
        case.compute_curvature_tensor()
        case.compute_force()
        case.update_vof(k, dt)
        case.compute_density(1)
        case.compute_mass_flux()
        case.predict_field(k, dt)
        dt_old, dt_new = substep_dts(dt, k)
        case.compute_potential(dt_old, dt_new)
        case.correct_velocity(dt_old, dt_new)
        case.update_pressure()
    return dt


class CaseSpec(C.Structure):
    """sbfs_case_t (include/sbfs.h)."""
    _fields_ = [("nx", C.c_size_t), ("ny", C.c_size_t), ("lx", C.c_double), ("ly", C.c_double),
                ("Re", C.c_double), ("Fr", C.c_double), ("denr", C.c_double), ("visr", C.c_double),
                ("We", C.c_double)]


class Ensemble:
    """Independent cases stepped together (sbfs_ensemble_*); `cases[k]` is a
    Case view of member k for set-up and read-back."""

    def __init__(self, backend: Backend, specs: list):
        self.be = backend
        lib = backend.lib
        lib.sbfs_ensemble_create.argtypes = [_P(CaseSpec), C.c_int, _P(C.c_void_p)]
        lib.sbfs_ensemble_case.argtypes = [C.c_void_p, C.c_int, _P(C.c_void_p), _P(C.c_void_p), _P(C.c_void_p)]
        lib.sbfs_ensemble_step.argtypes = [C.c_void_p, _P(C.c_double)]
        lib.sbfs_ensemble_autotune.argtypes = [C.c_void_p, _P(C.c_int)]
        lib.sbfs_ensemble_set_band.argtypes = [C.c_void_p, C.c_int]
        lib.sbfs_ensemble_destroy.argtypes = [C.c_void_p]
        arr = (CaseSpec * len(specs))(*[CaseSpec(s["nx"], s["ny"], s["lx"], s["ly"], s["Re"], s["Fr"], s["denr"],
                                                 s["visr"], s["We"]) for s in specs])
        self.handle = C.c_void_p()
        backend.check(lib.sbfs_ensemble_create(arr, len(specs), C.byref(self.handle)), "sbfs_ensemble_create")
        self.cases = []
        for k in range(len(specs)):
            d, f, i = C.c_void_p(), C.c_void_p(), C.c_void_p()
            backend.check(lib.sbfs_ensemble_case(self.handle, k, C.byref(d), C.byref(f), C.byref(i)),
                          "sbfs_ensemble_case")
            self.cases.append(Case.view(backend, d.value, f.value, i.value))
        self.lib = lib

    def step(self) -> np.ndarray:
        dts = np.zeros(len(self.cases))
        self.be.check(self.lib.sbfs_ensemble_step(self.handle, dts.ctypes.data_as(_P(C.c_double))),
                      "sbfs_ensemble_step")
        return dts

    def autotune(self) -> int:
        b = C.c_int(0)
        self.be.check(self.lib.sbfs_ensemble_autotune(self.handle, C.byref(b)), "sbfs_ensemble_autotune")
        return b.value

    def set_band(self, band: int) -> None:
        self.be.check(self.lib.sbfs_ensemble_set_band(self.handle, band), "sbfs_ensemble_set_band")

    def close(self) -> None:
        if self.handle:
            self.lib.sbfs_ensemble_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
