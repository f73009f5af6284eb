"""GPU parity of the plain and batched interpreters against the CPU oracle
(restatement of cl_cpu_eval.cpp:16-168, pinned to the compiled reference).

Tolerances (stated per result):
  * per-point predictions: |gpu - cpu32| <= 4 ulp(cpu32) + 2 |cpu32 - cpu64|
    (the oracle's own error against the FP64 evaluator bounds the
    conditioning of each point), non-finite classes equal;
  * MSE: relative error <= 1e-5, NaN/Inf classes equal.
"""
import numpy as np
import pytest

from oracle import oracle as O
from symbolic_regression_gpu_amd import (EVAL_BATCHED, EVAL_PLAIN, Population, SrError)
from symbolic_regression_gpu_amd.api import Tuning
from tests.helpers import (ALL_BINARY, ALL_UNARY, GRAMMAR_BINARY, GRAMMAR_UNARY, ill_conditioned, mse_close,
                           point_tolerance_ok, random_population, rel_err, to_api, uniform_data)

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
    c64 = O.predict64(tok, val, X, 3, N)
    ok = point_tolerance_ok(pred, c32, c64)
    _report("C1 predictions", pred.astype(np.float64), c32.astype(np.float64), c64)
    assert ok.all(), np.flatnonzero(~ok)[:10]
    m_ref, v_ref = O.mse32(pop, X, 3, N)
    for mode in (EVAL_PLAIN, EVAL_BATCHED):
        m, v, st = ctx.eval_mse(api, mode)
        assert mse_close(m, m_ref).all(), (mode, m, m_ref)
        assert (v == v_ref).all()
        assert st["logical_node_evals"] == int(pop.len.sum()) * N


@pytest.mark.parametrize("nv", [1, 3, 10, 20])
def test_population_mse_all_ops(ctx, nv):
    pop = random_population(120, 100 + nv, nv=nv, max_tokens=60, max_depth=7,
                            unary=ALL_UNARY, binary=ALL_BINARY, const_min=-3, const_max=3)
    N = 20_003
    X = uniform_data(nv, N, nv)
    X[nv] = np.sin(X[0]) * 2 + 1
    ctx.set_data(X)
    m_ref, v_ref = O.mse32(pop, X, nv, N, threads=8)
    m64 = O.mse64(pop, X, nv, N, threads=8)
    # Classify expressions: well-conditioned ones (the fp32 oracle is a
    # faithful answer at every point) get the strict MSE bound; the others are
    # checked point by point with the conditioning-aware rule.
    c32 = [O.predict32(*pop.expr(e), X, nv, N) for e in range(pop.E)]
    c64 = [O.predict64(*pop.expr(e), X, nv, N) for e in range(pop.E)]
    ill = np.array([ill_conditioned(a, b).sum() for a, b in zip(c32, c64)])
    well = ill == 0
    print(f"nv={nv}: {int((~well).sum())} of {pop.E} expressions have ill-conditioned points")
    api = to_api(pop)
    for mode in (EVAL_PLAIN, EVAL_BATCHED):
        m, v, _ = ctx.eval_mse(api, mode)
        close = mse_close(m, m_ref)
        bad = np.flatnonzero(well & ~(close & (v == v_ref)))
        assert len(bad) == 0, (mode, bad, m[bad], m_ref[bad])
        fin = well & np.isfinite(m) & np.isfinite(m64)
        print(f"nv={nv} mode={mode}: max rel err vs FP64 MSE {np.max(rel_err(m[fin], m64[fin])):.3e}"
              f" (oracle {np.max(rel_err(m_ref[fin], m64[fin])):.3e})")
    pred = ctx.eval_predictions(api)
    for e in range(pop.E):
        ok = point_tolerance_ok(pred[e], c32[e], c64[e])
        assert ok.all(), (e, np.flatnonzero(~ok)[:5])


def test_batched_tunings_agree(ctx):
    pop = random_population(64, 7, nv=3, max_tokens=60, max_depth=6,
                            unary=GRAMMAR_UNARY, binary=GRAMMAR_BINARY)
    N = 50_000
    X = uniform_data(3, N, 9)
    ctx.set_data(X)
    api = to_api(pop)
    m_ref, _ = O.mse32(pop, X, 3, N, threads=8)
    saved = ctx.tuning
    try:
        for dpt in (1, 2, 4, 8):
            for nq in (1, 8):
                for chunk in (512, 16384):
                    ctx.tuning = Tuning(dpt=dpt, chunk=chunk, n_queues=nq, blocks_per_cu=0)
                    m, _, _ = ctx.eval_mse(api, EVAL_BATCHED)
                    assert mse_close(m, m_ref).all(), (dpt, nq, chunk)
    finally:
        ctx.tuning = saved


@pytest.mark.parametrize("N", [1, 63, 64, 65, 1000, 4097])
def test_ragged_datapoint_counts(ctx, N):
    pop = random_population(9, 21, nv=2, max_tokens=20, max_depth=4)
    X = uniform_data(2, N, N)
    ctx.set_data(X)
    api = to_api(pop)
    m_ref, v_ref = O.mse32(pop, X, 2, N)
    for mode in (EVAL_PLAIN, EVAL_BATCHED):
        m, v, _ = ctx.eval_mse(api, mode)
        assert mse_close(m, m_ref).all()
        assert (v == v_ref).all()
    pred = ctx.eval_predictions(api)
    for e in range(pop.E):
        tok, val = pop.expr(e)
        c32 = O.predict32(tok, val, X, 2, N)
        c64 = O.predict64(tok, val, X, 2, N)
        assert point_tolerance_ok(pred[e], c32, c64).all()


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
