"""hipRTC backend on the host (no GPU): the generated module for seeded
populations compiles for gfx950, constants keep their exact fp32 bits, and
malformed expressions are rejected before code generation."""
import re

import numpy as np
import pytest

from symbolic_regression_gpu_amd import Population, SrError
from symbolic_regression_gpu_amd.api import jit_compile, jit_source
from tests.helpers import ALL_BINARY, ALL_UNARY, GRAMMAR_BINARY, GRAMMAR_UNARY, random_population, to_api


def test_source_structure_and_constants():
    pop = Population.from_exprs([([1, -1, 0], [0, 2.0, 0.1]), ([10, 8, 0, -1], [0, 0, -1.5, 0.0])])
    src = jit_source(pop, nv=3, dpt=2)
    assert "#define SR_NV 3" in src and "#define SR_DPT 2" in src
    # x[2] + 0.1f, with 0.1 as its fp32 bit pattern
    bits = np.float32(0.1).view(np.uint32)
    assert f"f_add(x[2], __int_as_float(0x{bits:08x}))" in src.replace("t0 = ", "")
    # sin(loose_div(-1.5, x[0]))
    b15 = np.float32(-1.5).view(np.uint32)
    assert f"f_loose_div(__int_as_float(0x{b15:08x}), x[0])" in src
    assert re.search(r"case 1: sr_chunk<SR_DPT>\(e1,", src)


def test_variable_beyond_nv_is_zero():
    pop = Population.from_exprs([([1, -1, -1], [0, 5.0, 0.0])])
    src = jit_source(pop, nv=2, dpt=4)
    assert "f_add(0.0f, x[0])" in src


@pytest.mark.parametrize("unary,binary,nv", [(GRAMMAR_UNARY, GRAMMAR_BINARY, 3), (ALL_UNARY, ALL_BINARY, 10)])
def test_generated_module_compiles(unary, binary, nv):
    pop = random_population(24, 77 + nv, nv=nv, max_tokens=60, max_depth=7, unary=unary, binary=binary)
    ms, nbytes = jit_compile(to_api(pop), nv, 4)
    assert nbytes > 1000 and ms > 0
    print(f"hipRTC compile of {pop.E} expressions ({int(pop.len.sum())} tokens): {ms:.0f} ms, {nbytes} B")


def test_malformed_rejected():
    for tok, val in [([1, 0], [0, 1.0]), ([-1], [-1.0]), ([-2], [0.0])]:
        with pytest.raises(SrError):
            jit_source(Population.from_exprs([(tok, val)]), nv=2)

