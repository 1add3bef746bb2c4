"""The reference's failure modes on the host side of the library, without a
GPU: each call returns non-zero and leaves the reason in sbfs_last_error.

* NPY headers (src/fileio.c:200-225): ndims, shape, dtype and storage
  order are checked, in that order, for every file domain_init reads.
* Uniform x grid (src/domain.c:77-96, :213-219): a non-uniform xf / xc is
  refused.
* Missing files and null arguments.
The environment-variable checks of fluid_init / interface_init
(src/config.c:14-42) run after the device allocation; their test is
tests/test_gpu_error_paths.py.
"""
import ctypes as C

import numpy as np
import pytest

from simplebubblyflowsolver_amd import abi, ic, solver


@pytest.fixture(scope="module")
def lib():
    lib = solver.load_library()
    lib.domain_init.argtypes = [C.c_char_p, C.POINTER(abi.DomainT)]
    lib.domain_init.restype = C.c_int
    lib.sbfs_last_error.restype = C.c_char_p
    return lib


def ic_dir(tmp_path, nx=16, ny=32):
    d = tmp_path / "ic"
    d.mkdir()
    xf, xc = ic.grid(nx, ny, 1.0, 2.0)[:2]
    np.save(d / "glsizes.npy", np.array([nx, ny], dtype="<u8"))
    np.save(d / "lengths.npy", np.array([1.0, 2.0]))
    np.save(d / "xf.npy", xf)
    np.save(d / "xc.npy", xc)
    return d


def init(lib, d):
    dom = abi.DomainT()
    rc = lib.domain_init(str(d).encode(), C.byref(dom))
    return rc, lib.sbfs_last_error().decode()


def test_valid_directory_loads(lib, tmp_path):
    d = ic_dir(tmp_path)
    rc, msg = init(lib, d)
    assert rc == 0, msg


@pytest.mark.parametrize("name,value,expect", [
    ("glsizes", np.array([[16, 32]], dtype="<u8"), "ndims: 1 expected, 2 obtained"),
    ("xf", np.linspace(0.0, 1.0, 18), "shape[0]: 17 expected, 18 obtained"),
    ("lengths", np.array([1.0, 2.0], dtype="<f4"), "dtype: '<f8' expected, '<f4' obtained"),
    ("glsizes", np.array([16, 32], dtype="<i8"), "dtype: '<u8' expected, '<i8' obtained"),
])
def test_npy_header_mismatch(lib, tmp_path, name, value, expect):
    d = ic_dir(tmp_path)
    np.save(d / f"{name}.npy", value)
    rc, msg = init(lib, d)
    assert rc != 0
    assert f"{name}.npy" in msg and expect in msg, msg


def test_npy_fortran_order_refused(lib, tmp_path):
    d = ic_dir(tmp_path)
    xc = np.load(d / "xc.npy")
    # numpy writes 1-D arrays in C order: write the header by hand
    hdr = b"{'descr': '<f8', 'fortran_order': True, 'shape': (%d,), }" % xc.size
    hdr += b" " * (64 - (10 + len(hdr) + 1) % 64) + b"\n"
    (d / "xc.npy").write_bytes(b"\x93NUMPY\x01\x00" + len(hdr).to_bytes(2, "little") + hdr + xc.tobytes())
    rc, msg = init(lib, d)
    assert rc != 0 and "is_fortran_order: false expected, true obtained" in msg, msg


def test_non_uniform_x_grid_refused(lib, tmp_path):
    nx = 16
    d = ic_dir(tmp_path, nx)
    xf = np.linspace(0.0, 1.0, nx + 1) ** 1.2
    xc = np.concatenate([[0.0], 0.5 * (xf[:-1] + xf[1:]), [1.0]])
    np.save(d / "xf.npy", xf)
    np.save(d / "xc.npy", xc)
    rc, msg = init(lib, d)
    assert rc != 0 and "x grid is not uniform" in msg, msg


def test_missing_file_and_null_arguments(lib, tmp_path):
    d = ic_dir(tmp_path)
    (d / "lengths.npy").unlink()
    rc, msg = init(lib, d)
    assert rc != 0 and "lengths.npy" in msg, msg
    dom = abi.DomainT()
    assert lib.domain_init(None, C.byref(dom)) != 0
    assert "null argument" in lib.sbfs_last_error().decode()


def test_ensemble_refuses_fields_of_4_gib(lib):
    """The FP64 stage kernels address a field by a 32-bit byte offset
    (simplebubblyflowsolver_amd/csrc/march.hip), so a case whose vof field
    reaches 4 GiB is refused before anything is allocated."""
    lib.sbfs_ensemble_create.argtypes = [C.POINTER(solver.CaseSpec), C.c_int, C.POINTER(C.c_void_p)]
    case = solver.CaseSpec(32768, 16384, 1.0, 0.5, 35.0, 1.0, 1e-3, 1e-2, 10.0)
    handle = C.c_void_p()
    assert lib.sbfs_ensemble_create(C.byref(case), 1, C.byref(handle)) != 0
    assert "4 GiB" in lib.sbfs_last_error().decode()
