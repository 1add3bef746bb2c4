"""Mutation check of the oracle's pins.

Plausible mistakes are planted in a copy of oracle/oracle.c (on the CPU,
never in the product), the copy is built with the oracle's own flags, and
the closed-form pins of tests/test_oracle_pins_flow.py are run against it.
Each mistake must make at least the pin named for it fail; the unmodified
source must pass all of them (test_oracle_pins_flow.py).  The mistakes:

  M1  sign of the ux x-advection term (reference predict/ux.c:29-61)
  M2  THINC transverse Gauss points +-1/2 instead of +-1/(2 sqrt 3)
      (interface/common.c:4-11)
  M3  visux and visuy exchanged in the shear stress (predict/txy.c:29-40)
  M4  the upper mass flux of the uy y-advection taken one row low
      (predict/uy.c:63-90)
  M5  the previous stage's uy source weighted by alpha instead of beta
      (predict/uy.c:214-251)
  M6  sign of the pressure gradient in the ux predictor (predict/ux.c:146-162)
"""
import pytest

import oracle
import test_oracle_pins_flow as pins

MUTANTS = {
    "M1": ("0.5 * UXa(i + 1, j);\n          s -= 1. / JDXF(i) * (-mm * um + mp * up);",
           "0.5 * UXa(i + 1, j);\n          s += 1. / JDXF(i) * (-mm * um + mp * up);",
           "ux_advection_in_x"),
    "M2": ("{-0.2886751345948129, +0.2886751345948129}", "{-0.5, +0.5}",
           "thinc_oblique_face_average"),
    "M3": ("TXYa(i, j) = vx * lxy + vy * lyx;", "TXYa(i, j) = vy * lxy + vx * lyx;",
           "shear_stress_uses_x_face_viscosity"),
    "M4": ("FLUXYa(i, j) + 0.5 * JDXC(i) / hy * FLUXYa(i, j + 1);",
           "FLUXYa(i, j) + 0.5 * JDXC(i) / hy * FLUXYa(i, j - 1);",
           "uy_advection_and_diffusion_in_y"),
    "M5": ("cb * dt * SRCUYa(sb, i, j)", "ca * dt * SRCUYa(sb, i, j)",
           "decaying_shear_wave"),
    "M6": ("SRCUXa(sg, i, j) -= 1. / HXXF(i)", "SRCUXa(sg, i, j) += 1. / HXXF(i)",
           "stratified_column_at_rest"),
}


def caught(build, check) -> bool:
    try:
        check(build)
    except Exception:
        return True
    return False


@pytest.mark.parametrize("mutant", sorted(MUTANTS))
def test_planted_mistake_fails_its_pin(tmp_path, mutant):
    old, new, pin = MUTANTS[mutant]
    src = (oracle.HERE / "oracle.c").read_text()
    assert src.count(old) == 1, f"{mutant}: anchor not unique in oracle.c"
    lib = oracle.compile_library(tmp_path / f"liboracle_{mutant}.so", src.replace(old, new))
    build = oracle.Build(lib)
    failing = sorted(name for name, check in pins.CHECKS.items() if caught(build, check))
    print(f"{mutant}: fails {failing}")
    assert pin in failing, (mutant, failing)
