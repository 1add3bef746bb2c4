"""The C-ABI library loads and exports every entry point the public
headers declare (no device calls)."""
import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
LIB = ROOT / "simplebubblyflowsolver_amd" / "libsbfs.so"


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        txt = h.read_text()
        for m in re.finditer(r"extern\s+(?:const\s+)?[A-Za-z_]+\s*\**\s*([A-Za-z_][A-Za-z0-9_]*)\s*\(", txt):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_the_reference_entry_points():
    names = set(declared_functions())
    reference = {
        "integrate", "decide_dt", "domain_init", "domain_save",
        "fluid_init", "fluid_save", "fluid_compute_density", "fluid_compute_viscosity",
        "fluid_predict_field", "fluid_compute_potential", "fluid_correct_velocity",
        "fluid_update_pressure", "fluid_update_boundaries_ux", "fluid_update_boundaries_uy",
        "fluid_update_boundaries_p", "fluid_update_boundaries_psi",
        "interface_init", "interface_save", "interface_compute_curvature_tensor",
        "interface_compute_force", "interface_compute_mass_flux", "interface_update_vof",
        "interface_update_boundaries_vof",
    }
    assert reference <= names, reference - names


def test_library_exports_every_declared_symbol():
    if not LIB.exists():
        pytest.fail(f"{LIB} is not built")
    lib = C.CDLL(str(LIB))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts_match_the_headers():
    """Offsets of the ctypes mirrors equal those of the C structs, checked
    with a tiny C program compiled against include/."""
    import subprocess
    import tempfile
    from simplebubblyflowsolver_amd import abi
    src = r"""
#include <stdio.h>
#include <stddef.h>
#include "domain.h"
#include "array.h"
#include "fluid.h"
#include "interface.h"
int main(void){
  printf("%zu %zu %zu %zu\n", sizeof(array_t), sizeof(domain_t), sizeof(fluid_t), sizeof(interface_t));
  printf("%zu %zu %zu\n", offsetof(domain_t, hy), offsetof(fluid_t, refden), offsetof(interface_t, We));
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        p = Path(d)
        (p / "t.c").write_text(src)
        subprocess.run(["gcc", "-std=c99", f"-I{ROOT / 'include'}", str(p / "t.c"), "-o", str(p / "t")], check=True)
        out = subprocess.run([str(p / "t")], capture_output=True, text=True, check=True).stdout.split()
    sizes = [int(v) for v in out]
    assert sizes[:4] == [C.sizeof(abi.ArrayT), C.sizeof(abi.DomainT), C.sizeof(abi.FluidT), C.sizeof(abi.InterfaceT)]
    assert sizes[4] == abi.DomainT.hy.offset
    assert sizes[5] == abi.FluidT.refden.offset
    assert sizes[6] == abi.InterfaceT.We.offset
