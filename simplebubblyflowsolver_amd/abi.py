"""ctypes mirrors of the C-ABI structs (include/array.h, domain.h, fluid.h,
interface.h) and layout helpers.

The struct layouts are those of the reference's array_t / domain_t /
fluid_t / interface_t (reference include/array.h:10-23, include/domain.h:20-30,
include/fluid.h:27-45, include/interface.h:13-24).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

NDIMS = 2


class ArrayT(C.Structure):
    _fields_ = [
        ("size", C.c_size_t),
        ("nadds", C.POINTER(C.c_int * 2)),
        ("datasize", C.c_size_t),
        ("data", C.c_void_p),
    ]


class DomainT(C.Structure):
    _fields_ = [
        ("info", C.c_void_p),
        ("glsizes", C.c_size_t * NDIMS),
        ("mysizes", C.c_size_t * NDIMS),
        ("offsets", C.c_size_t * NDIMS),
        ("lengths", C.c_double * NDIMS),
        ("xf", C.POINTER(C.c_double)),
        ("xc", C.POINTER(C.c_double)),
        ("jdxf", C.POINTER(C.c_double)),
        ("jdxc", C.POINTER(C.c_double)),
        ("hxxf", C.POINTER(C.c_double)),
        ("hxxc", C.POINTER(C.c_double)),
        ("hy", C.c_double),
    ]


class FluidT(C.Structure):
    _fields_ = [
        ("ux", ArrayT),
        ("uy", ArrayT),
        ("p", ArrayT),
        ("psi", ArrayT * 2),
        ("den", ArrayT * 2),
        ("visux", ArrayT),
        ("visuy", ArrayT),
        ("txx", ArrayT),
        ("txy", ArrayT),
        ("tyy", ArrayT),
        ("srcux", ArrayT * 3),
        ("srcuy", ArrayT * 3),
        ("Re", C.c_double),
        ("Fr", C.c_double),
        ("denr", C.c_double),
        ("visr", C.c_double),
        ("refden", C.c_double),
    ]


class InterfaceT(C.Structure):
    _fields_ = [
        ("vof", ArrayT),
        ("ifrcx", ArrayT),
        ("ifrcy", ArrayT),
        ("dvof", ArrayT),
        ("normal", ArrayT),
        ("curv", ArrayT),
        ("fluxx", ArrayT),
        ("fluxy", ArrayT),
        ("src", ArrayT * 2),
        ("We", C.c_double),
    ]


# Field names used by the Python helpers -> (struct, member, index or None).
FIELDS = {
    "ux": ("fluid", "ux", None),
    "uy": ("fluid", "uy", None),
    "p": ("fluid", "p", None),
    "psi0": ("fluid", "psi", 0),
    "psi1": ("fluid", "psi", 1),
    "den0": ("fluid", "den", 0),
    "den1": ("fluid", "den", 1),
    "visux": ("fluid", "visux", None),
    "visuy": ("fluid", "visuy", None),
    "txx": ("fluid", "txx", None),
    "txy": ("fluid", "txy", None),
    "tyy": ("fluid", "tyy", None),
    "srcux_a": ("fluid", "srcux", 0),
    "srcux_b": ("fluid", "srcux", 1),
    "srcux_g": ("fluid", "srcux", 2),
    "srcuy_a": ("fluid", "srcuy", 0),
    "srcuy_b": ("fluid", "srcuy", 1),
    "srcuy_g": ("fluid", "srcuy", 2),
    "vof": ("interface", "vof", None),
    "ifrcx": ("interface", "ifrcx", None),
    "ifrcy": ("interface", "ifrcy", None),
    "dvof": ("interface", "dvof", None),
    "normal": ("interface", "normal", None),
    "curv": ("interface", "curv", None),
    "fluxx": ("interface", "fluxx", None),
    "fluxy": ("interface", "fluxy", None),
    "src_a": ("interface", "src", 0),
    "src_b": ("interface", "src", 1),
}

# ghost / halo widths {x: (lo, hi), y: (lo, hi)} (reference include/array_macros/)
NADDS = {
    "ux": ((0, 1), (1, 1)),
    "uy": ((1, 1), (1, 1)),
    "p": ((1, 1), (1, 1)),
    "psi0": ((1, 1), (1, 1)),
    "psi1": ((1, 1), (1, 1)),
    "den0": ((1, 1), (1, 1)),
    "den1": ((1, 1), (1, 1)),
    "visux": ((0, 1), (1, 1)),
    "visuy": ((1, 1), (1, 1)),
    "txx": ((0, 0), (1, 1)),
    "txy": ((0, 1), (1, 1)),
    "tyy": ((0, 0), (1, 1)),
    "srcux_a": ((-1, 0), (0, 0)),
    "srcux_b": ((-1, 0), (0, 0)),
    "srcux_g": ((-1, 0), (0, 0)),
    "srcuy_a": ((0, 0), (0, 0)),
    "srcuy_b": ((0, 0), (0, 0)),
    "srcuy_g": ((0, 0), (0, 0)),
    "vof": ((1, 1), (2, 2)),
    "ifrcx": ((-1, 0), (0, 0)),
    "ifrcy": ((0, 0), (0, 0)),
    "dvof": ((0, 1), (1, 2)),
    "normal": ((0, 0), (1, 1)),
    "curv": ((1, 1), (1, 1)),
    "fluxx": ((0, 1), (1, 1)),
    "fluxy": ((1, 1), (1, 1)),
    "src_a": ((0, 0), (0, 0)),
    "src_b": ((0, 0), (0, 0)),
}

COMPONENTS = {"dvof": 2, "normal": 3}


def field_shape(name: str, nx: int, ny: int) -> tuple:
    """numpy shape (rows=y, cols=x[, comps]) of a whole buffer incl. ghosts."""
    (xl, xh), (yl, yh) = NADDS[name]
    shape = (ny + yl + yh, nx + xl + xh)
    comps = COMPONENTS.get(name)
    return shape + ((comps,) if comps else ())


def field_origin(name: str) -> tuple:
    """(row, col) of the interior element (i=1, j=1) in the whole buffer."""
    (xl, _), (yl, _) = NADDS[name]
    return yl, xl


def interior(name: str, arr: np.ndarray, nx: int, ny: int) -> np.ndarray:
    """View of the interior cells (j=1..ny, i=1..nx) of a whole buffer."""
    r0, c0 = field_origin(name)
    return arr[r0:r0 + ny, c0:c0 + nx]


def get_array(fluid: FluidT, interface: InterfaceT, name: str) -> ArrayT:
    owner, member, index = FIELDS[name]
    obj = fluid if owner == "fluid" else interface
    a = getattr(obj, member)
    return a[index] if index is not None else a
