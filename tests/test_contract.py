"""The opt-in contract form (spec key ``contract``): FMA contraction across the
instruction list, division by a constant as the product with its reciprocal,
and division by a computed value through the hardware reciprocal refined by
Newton steps (``fmr_div_contract``).  It is not bit-exact, so it is checked
against the plain FP64 path -- which the parity tests hold bit-exact with the
oracle -- to the FP64 tolerance of 1e-12 relative max-norm per field, at the
stated steps:

* Euler HLLC 256^3, 200 steps (every division of the flux is by a state
  value: contraction and the refined reciprocal change the arithmetic);
* Gray-Scott 64^2, 300 steps.

Gray-Scott 2048^2 over its full 10^4 steps does not meet 1e-12 in this form
(the run amplifies rounding-size differences; measured 1e-10 relative
max-norm): that library stays a measured option, not the headline path.
"""
import numpy as np
import pytest

from conftest import hip_lib
from formura_amd import inputs

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _run(name, steps, seed_fn):
    h = hip_lib(name)
    h.init()
    h.set_state(seed_fn(h))
    h.set_resident(True)
    h.upload()
    h.forward_steps(steps)
    h.download()
    out = {f: h.field(f).copy() for f in h.fields}
    h.set_resident(False)
    h.finalize()
    return out


@pytest.mark.parametrize("plain,contract,steps,seed", [
    ("euler3d_256", "euler3d_256_contract", 200, lambda h: inputs.euler_state(1, h.grid)),
    ("gs_64", "gs_64_contract", 300,
     lambda h: inputs.grayscott_perturb(h.field("u").copy(), h.field("v").copy(), 7)),
])
def test_contract_form_within_fp64_tolerance(plain, contract, steps, seed):
    print("%s vs %s: %d steps" % (contract, plain, steps), flush=True)
    a = _run(plain, steps, seed)
    b = _run(contract, steps, seed)
    differs = False
    for f in a:
        err = float(np.max(np.abs(a[f] - b[f])) / np.max(np.abs(a[f])))
        print("  %s: relative max-norm %.3g" % (f, err), flush=True)
        assert err <= TOL, (f, err)
        differs |= err > 0
    assert differs  # the form does change the arithmetic
