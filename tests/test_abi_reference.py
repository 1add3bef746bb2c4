"""The drop-in boundary against the reference's own headers (read as text
where the reference tree is present; skipped elsewhere): the members of
array_t, domain_t, fluid_t and interface_t in the same order with the same
types and extents, and every entry point the reference's solver headers
declare with the same parameter list.  Comments, whitespace and the
spelling of `restrict` are normalised away; nothing of the reference is
compiled or kept.  (tests/test_abi.py checks the compiled layouts and the
library's exports against these headers.)"""
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
REF = Path("/root/reference/include")
STRUCTS = {"array.h": "array_t", "domain.h": "domain_t", "fluid.h": "fluid_t", "interface.h": "interface_t"}
SOLVER_HEADERS = ("fluid_solver.h", "interface_solver.h", "integrate.h", "decide_dt.h")

pytestmark = pytest.mark.skipif(not REF.is_dir(), reason="the reference tree is not present on this machine")


def clean(text: str) -> str:
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    text = text.replace("SBFS_RESTRICT", "restrict")
    return text


def norm(s: str) -> str:
    s = re.sub(r"\s+", " ", s).strip()
    s = re.sub(r"\s*([*\[\](),;])\s*", r"\1", s)
    return s


def struct_members(text: str, name: str) -> list:
    for m in re.finditer(r"typedef\s+struct\s*\{(.*?)\}\s*(\w+)\s*;", clean(text), flags=re.S):
        if m.group(2) == name:
            return [norm(d) for d in m.group(1).split(";") if d.strip()]
    raise AssertionError(f"{name} not found")


def prototypes(text: str) -> dict:
    out = {}
    for m in re.finditer(r"extern\s+int\s+(\w+)\s*\((.*?)\)\s*;", clean(text), flags=re.S):
        out[m.group(1)] = [norm(p) for p in m.group(2).split(",")]
    return out


@pytest.mark.parametrize("header,name", sorted(STRUCTS.items()))
def test_struct_members_match_the_reference(header, name):
    ref = struct_members((REF / header).read_text(), name)
    ours = struct_members((ROOT / "include" / header).read_text(), name)
    assert ours == ref


def test_entry_points_match_the_reference():
    ref, ours = {}, {}
    for h in SOLVER_HEADERS:
        ref.update(prototypes((REF / h).read_text()))
        ours.update(prototypes((ROOT / "include" / h).read_text()))
    assert ref, "no prototypes parsed from the reference headers"
    missing = sorted(set(ref) - set(ours))
    assert not missing, missing
    for name, params in ref.items():
        assert ours[name] == params, (name, ours[name], params)
