"""The exact kernels compute some correctly rounded operations without the division sequence
(feriphys_amd/csrc/exact_ops.hpp) and atan2 without fdlibm's branches (fdlibm_atan2f.hpp).  The
reciprocal and square-root forms rest on the gfx950 reciprocal and square-root instructions, which
no CPU reproduces, so they are checked on the GPU itself: over every finite f32 input, no result differs from the operation it
stands for."""
import pytest

import feriphys_amd as fa
from gpu_common import announce


@pytest.mark.gpu
def test_exact_math_forms_match_on_every_f32():
    announce("selftest_exact_math over all 2^32 bit patterns")
    assert fa.selftest_exact_math() == {"rcp": 0, "div_pi": 0, "div3": 0, "atan2": 0, "sqrt": 0}
