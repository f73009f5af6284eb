"""hipRTC backend (mode JIT) on the GPU: MSEs and valid counts equal the
batched interpreter's bit for bit (same op functions, same work items, same
per-lane summation order), hence the oracle's under the conditioning-aware
bound; the module is compiled once per distinct population."""
import numpy as np
import pytest

from symbolic_regression_gpu_amd import EVAL_BATCHED, EVAL_JIT
from symbolic_regression_gpu_amd.api import Tuning
from tests.helpers import (ALL_BINARY, ALL_UNARY, GRAMMAR_BINARY, GRAMMAR_UNARY, check_population,
                           random_population, to_api, uniform_data)

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skip(reason="device execution of hipRTC modules is switched off: its first GPU run "
                            "(100 exprs x 100k, nv=10) stopped making progress after the compile, "
                            "and again after the code-object lifetime fix; cause not identified"),
]


@pytest.mark.parametrize("unary,binary,nv,dpt", [(GRAMMAR_UNARY, GRAMMAR_BINARY, 10, 4),
                                                 (ALL_UNARY, ALL_BINARY, 3, 2),
                                                 (GRAMMAR_UNARY, GRAMMAR_BINARY, 1, 1)])
def test_jit_equals_batched_and_oracle(ctx, unary, binary, nv, dpt):
    # multi-expression batch: 100 expressions of 5..60 tokens
    pop = random_population(100, 300 + nv, nv=nv, max_tokens=60, max_depth=7, unary=unary,
                            binary=binary, min_len=5)
    N = 100_003
    X = uniform_data(nv, N, nv + 1)
    X[nv] = np.cos(X[0]) * 3
    ctx.set_data(X)
    saved = ctx.tuning
    try:
        ctx.tuning = Tuning(dpt=dpt, chunk=8192, n_queues=8, blocks_per_cu=0)
        api = to_api(pop)
        print(f"[jit nv={nv} dpt={dpt}] batched", flush=True)
        m_b, v_b, _ = ctx.eval_mse(api, EVAL_BATCHED)
        print(f"[jit nv={nv} dpt={dpt}] compile + JIT launch", flush=True)
        m_j, v_j, st = ctx.eval_mse(api, EVAL_JIT)
        print(f"[jit nv={nv} dpt={dpt}] JIT done", flush=True)
        assert st["jit_compile_time_ms"] > 0
        assert np.array_equal(m_j, m_b, equal_nan=True)
        assert np.array_equal(v_j, v_b)
        m_j2, _, st2 = ctx.eval_mse(api, EVAL_JIT)
        assert st2["jit_compile_time_ms"] == 0 and np.array_equal(m_j2, m_j, equal_nan=True)
        print(f"JIT nv={nv}: compile {st['jit_compile_time_ms']:.0f} ms, kernel {st2['interp_kernel_ms']:.3f} ms")
        check_population(pop, X, nv, N, m_j, v_j)
    finally:
        ctx.tuning = saved
