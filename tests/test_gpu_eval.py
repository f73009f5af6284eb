"""GPU parity of the plain and batched interpreters against the CPU oracle
(restatement of cl_cpu_eval.cpp:16-168, pinned to the compiled reference).

Tolerances (stated per result, see tests/helpers.py):
  * per point: |gpu - v64| <= 2 max(bound, |cpu32 - v64|) + 2 ulp, where v64 is
    the FP64 value and bound the first-order error bound of an fp32
    evaluation (oracle sro_predict64_bound), classes equal to the oracle's;
    points whose bound is infinite (a protected branch may flip) are exempt;
  * MSE: relative error <= 1e-5 against the oracle and equal valid counts
    for every expression whose MSE error bound is <= 1e-6; every MSE within
    twice its bound of the FP64 MSE.
"""
import numpy as np
import pytest

from oracle import oracle as O
from symbolic_regression_gpu_amd import (EVAL_BATCHED, EVAL_PLAIN, Population, SrError)
from symbolic_regression_gpu_amd.api import Tuning
from tests.helpers import (ALL_BINARY, ALL_UNARY, GRAMMAR_BINARY, GRAMMAR_UNARY, check_population,
                           mse_close, point_parity, random_population, rel_err, to_api,
                           uniform_data)

pytestmark = pytest.mark.gpu


def _report(name, gpu, c32, c64):
    fin = np.isfinite(gpu) & np.isfinite(c64)
    if fin.any():
        e_gpu = rel_err(gpu[fin], c64[fin])
        e_cpu = rel_err(c32[fin], c64[fin])
        print(f"[{name}] max rel err vs FP64: gpu {np.nanmax(e_gpu):.3e}  cpu32 {np.nanmax(e_cpu):.3e}")


def test_single_expression_predictions_c1_shape(ctx):
    # One depth<=6, <=31-token tree over 3 vars ~ U[1,5], consts U[-2,2], N=100k.
    pop = random_population(1, 11, nv=3, max_tokens=31, max_depth=6, unary=GRAMMAR_UNARY,
                            binary=GRAMMAR_BINARY, min_len=9)
    target = random_population(1, 12, nv=3, max_tokens=31, max_depth=6, unary=GRAMMAR_UNARY,
                               binary=GRAMMAR_BINARY, min_len=9)
    N = 100_000
    X = uniform_data(3, N, 5)
    t_tok, t_val = target.expr(0)
    X[3] = O.predict32(t_tok, t_val, X, 3, N)
    ctx.set_data(X)
    api = to_api(pop)
    pred = ctx.eval_predictions(api)[0]
    tok, val = pop.expr(0)
    c32 = O.predict32(tok, val, X, 3, N)
    c64, b64 = O.predict64_bound(tok, val, X, 3, N)
    ok = point_parity(pred, c32, c64, b64)
    _report("C1 predictions", pred.astype(np.float64), c32.astype(np.float64), c64)
    assert ok.all(), np.flatnonzero(~ok)[:10]
    for mode in (EVAL_PLAIN, EVAL_BATCHED):
        m, v, st = ctx.eval_mse(api, mode)
        check_population(pop, X, 3, N, m, v)
        assert st["logical_node_evals"] == int(pop.len.sum()) * N


# variable classes of the batched kernel at its default shape: nv <= 3 -> 3
# registers, 4 -> 4, 5..16 -> 16, 17..32 -> 32
@pytest.mark.parametrize("nv", [1, 3, 4, 10, 20, 32])
def test_population_mse_all_ops(ctx, nv):
    pop = random_population(120, 100 + nv, nv=nv, max_tokens=60, max_depth=7,
                            unary=ALL_UNARY, binary=ALL_BINARY, const_min=-3, const_max=3)
    N = 20_003
    X = uniform_data(nv, N, nv)
    X[nv] = np.sin(X[0]) * 2 + 1
    ctx.set_data(X)
    api = to_api(pop)
    pred = ctx.eval_predictions(api)
    for mode in (EVAL_PLAIN, EVAL_BATCHED):
        m, v, _ = ctx.eval_mse(api, mode)
        st = check_population(pop, X, nv, N, m, v, pred_gpu=pred if mode == EVAL_BATCHED else None,
                              name=f"all ops nv={nv} mode={mode}")
        assert st["strict"] >= pop.E // 8, st


def test_batched_tunings_agree(ctx):
    pop = random_population(64, 7, nv=3, max_tokens=60, max_depth=6,
                            unary=GRAMMAR_UNARY, binary=GRAMMAR_BINARY)
    N = 50_000
    X = uniform_data(3, N, 9)
    ctx.set_data(X)
    api = to_api(pop)
    base, vbase, _ = ctx.eval_mse(api, EVAL_BATCHED)
    check_population(pop, X, 3, N, base, vbase)
    saved = ctx.tuning
    by_shape = {}
    try:
        # dpt 0: the vector interpreter
        for dpt in (0, 1, 2, 3, 4):
            for nq in (1, 8):
                for chunk in (512, 16384):
                    ctx.tuning = Tuning(dpt=max(dpt, 1), chunk=chunk, n_queues=nq, blocks_per_cu=0,
                                        kernel=int(dpt == 0))
                    m, v, _ = ctx.eval_mse(api, EVAL_BATCHED)
                    # same per-point values; only the fp64 partial-sum order moves
                    assert np.array_equal(v, vbase), (dpt, nq, chunk)
                    assert mse_close(m, base, 1e-12).all(), (dpt, nq, chunk)
                    # chunk boundaries and per-lane summation order do not
                    # depend on dpt: equal chunk and queues give equal bits
                    ref = by_shape.setdefault((nq, chunk), m)
                    assert np.array_equal(m, ref, equal_nan=True), (dpt, nq, chunk)
    finally:
        ctx.tuning = saved


@pytest.mark.parametrize("N", [1, 63, 64, 65, 1000, 4097])
def test_ragged_datapoint_counts(ctx, N):
    pop = random_population(9, 21, nv=2, max_tokens=20, max_depth=4)
    X = uniform_data(2, N, N)
    ctx.set_data(X)
    api = to_api(pop)
    pred = ctx.eval_predictions(api)
    for mode in (EVAL_PLAIN, EVAL_BATCHED):
        m, v, _ = ctx.eval_mse(api, mode)
        check_population(pop, X, 2, N, m, v, pred_gpu=pred)


def test_edge_expressions(ctx):
    N = 777
    X = uniform_data(3, N, 3)
    ctx.set_data(X)
    exprs = [
        ([0], [2.5]),                       # constant
        ([-1], [2.0]),                      # variable
        ([-1], [9.0]),                      # variable index >= nv -> 0
        ([4, -1, 0], [0, 0, 0.0]),          # DIV by zero -> NaN everywhere -> MSE NaN
        ([16, 0], [0, 200.0]),              # exp overflow -> inf -> excluded -> NaN
        ([29, 0], [0, 11.0]),               # clamped exp -> 1e9
        ([8, -1, 0], [0, 1, 0.0]),          # LOOSE_DIV by 0 -> x/1e-9
        ([22, 0], [0, 0.0]),                # LOOSE_LOG(0) -> -1e9
        ([9, 0, 0], [0, 0.0, 0.0]),         # LOOSE_POW(0,0) -> 0
        ([18, 0], [0, 0.0]),                # INV(0) -> +inf -> excluded
        ([0], [1e20]),                      # finite pred, squared error overflows -> inf MSE
        ([3, 0, 0], [0, 1e20, 1e20]),       # product overflows -> inf pred -> excluded
    ]
    pop = O.Population(exprs)
    api = to_api(pop)
    m_ref, v_ref = O.mse32(pop, X, 3, N)
    for mode in (EVAL_PLAIN, EVAL_BATCHED):
        m, v, _ = ctx.eval_mse(api, mode)
        assert mse_close(m, m_ref).all(), (mode, m, m_ref)
        assert (v == v_ref).all()
    assert np.isnan(m_ref[3]) and np.isnan(m_ref[4]) and np.isnan(m_ref[9])


def test_malformed_population_is_an_error(ctx):
    X = uniform_data(1, 100, 1)
    ctx.set_data(X)
    bad = Population([1, 0], [0, 1.0], [0], [2])
    for mode in (EVAL_PLAIN, EVAL_BATCHED):
        with pytest.raises(SrError) as ei:
            ctx.eval_mse(bad, mode)
        assert ei.value.code == -3


def test_autotune_picks_a_valid_shape_and_keeps_results(ctx):
    pop = random_population(200, 17, nv=3, max_tokens=80, max_depth=7,
                            unary=GRAMMAR_UNARY, binary=GRAMMAR_BINARY)
    N = 65_537
    X = uniform_data(3, N, 19)
    ctx.set_data(X)
    api = to_api(pop)
    saved = ctx.tuning
    try:
        base, vbase, _ = ctx.eval_mse(api, EVAL_BATCHED)
        best, ms = ctx.autotune(api, max_sample=64)
        assert best.dpt in (1, 2, 3, 4) and best.n_queues in (1, 8) and best.chunk % 768 == 0
        assert ms > 0 and ctx.tuning == best
        m, v, _ = ctx.eval_mse(api, EVAL_BATCHED)
        assert np.array_equal(v, vbase) and mse_close(m, base, 1e-12).all()
        again, ms2 = ctx.autotune(api, max_sample=64)   # remembered per shape
        assert again == best and ms2 == 0
        print(f"autotune: {best} ({ms:.3f} ms on the sample)")
    finally:
        ctx.tuning = saved


def _deep(k, ops=(1, 2, 3, 4), var=0):
    """A tree whose bytecode needs k + 1 stack slots: op(T(k-1), T(k-1))."""
    if k == 0:
        return [ops[0], -1, -1], [0.0, float(var), float(var + 1)]
    lt, lv = _deep(k - 1, ops[1:] + ops[:1], var)
    rt, rv = _deep(k - 1, ops[2:] + ops[:2], var + 1)
    return [ops[k % len(ops)]] + lt + rt, [0.0] + lv + rv


@pytest.mark.parametrize("nv,N", [(1, 20_003), (3, 100_000), (4, 4097), (10, 65_536), (32, 1000)])
def test_vector_interpreter_equals_lane(ctx, nv, N):
    """The vector interpreter (384 datapoints per instruction, stack in LDS
    with deeper slots spilled) gives the lane interpreter's MSEs bit for bit:
    same op functions, same operand order, same per-lane summation order."""
    pop = random_population(150, 700 + nv, nv=nv, max_tokens=60, max_depth=7, unary=ALL_UNARY,
                            binary=ALL_BINARY, const_min=-3, const_max=3)
    exprs = [pop.expr(e) for e in range(pop.E)]
    for k in (3, 5, 6, 7):  # stack needs 4, 6, 7, 8: slots 4..7 live in the spill area
        t, v = _deep(k, var=0)
        exprs.append((np.array(t, np.int32), np.array([x % max(nv, 1) if tt == -1 else x
                                                      for tt, x in zip(t, v)], np.float64)))
    pop = O.Population(exprs)
    X = uniform_data(nv, N, nv + 3)
    X[nv] = np.cos(X[0]) + 0.5
    ctx.set_data(X)
    api = to_api(pop)
    saved = ctx.tuning
    try:
        ctx.tuning = Tuning(dpt=3, chunk=3072, n_queues=8, blocks_per_cu=0, kernel=0)
        m_lane, v_lane, _ = ctx.eval_mse(api, EVAL_BATCHED)
        ctx.tuning = Tuning(dpt=3, chunk=3072, n_queues=8, blocks_per_cu=0, kernel=1)
        m_vec, v_vec, _ = ctx.eval_mse(api, EVAL_BATCHED)
    finally:
        ctx.tuning = saved
    assert np.array_equal(m_vec.view(np.uint32), m_lane.view(np.uint32))
    assert np.array_equal(v_vec, v_lane)
    check_population(pop, X, nv, N, m_vec, v_vec, name=f"vector nv={nv}")
