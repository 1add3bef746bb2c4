"""The closed-form pins of tests/test_oracle_pins_flow.py run on the HIP
library itself: advection, variable-viscosity diffusion and the shear
stress's viscosity placement (stage entry points, staged kernels), the
THINC face average of an oblique interface, a stratified column at rest and
a decaying shear wave (integrate, the fused graph-replayed step), in FP64
and in mixed precision where the check is a whole step.  These hold the
library to the mathematics directly, beside its bitwise match with the
oracle."""
import pytest

import test_oracle_pins_flow as pins
from simplebubblyflowsolver_amd import solver

pytestmark = pytest.mark.gpu


class Product:
    """`new_case` on the HIP library, for the `check_*` functions."""

    def __init__(self, be):
        self.be = be

    def new_case(self, nx, ny, lx=1.0, ly=2.0, Re=35.0, Fr=1.0, denr=1e-3, visr=1e-2, We=10.0):
        return solver.Case(self.be, nx, ny, lx, ly, Re, Fr, denr, visr, We)


@pytest.mark.parametrize("name", list(pins.CHECKS))
def test_closed_form_on_gpu(product, name):
    pins.CHECKS[name](Product(product))


def test_whole_step_closed_forms_mixed_precision(product):
    """The FP32 stencils leave both whole-step closed forms within their
    FP64 bounds: the column stays at rest (the divergence constraint is
    enforced in FP64) and the shear wave is a linear FP64 update of FP32
    sources, held here to 1e-6 of its amplitude."""
    lib = product.lib
    assert lib.sbfs_set_precision(1) == 0
    try:
        pins.check_stratified_column_at_rest(Product(product))
        pins.check_decaying_shear_wave(Product(product), rtol=1e-6)
    finally:
        assert lib.sbfs_set_precision(0) == 0
