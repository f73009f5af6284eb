"""Trace replay (evaluate_evolution_benchmark, utils.cpp:1275-1393) and the
SR system's per-generation CSV on the GPU.

* a trace written from the device GA replays to exactly the fitness vectors
  the GA computed (batched and both cache policies bit for bit, plain within
  the MSE bar);
* the reference's example digest (two expressions, two data files: the
  non-shared path) replays to the oracle's MSEs.
"""
import os
import shutil

import numpy as np
import pytest

from oracle import oracle as O
from symbolic_regression_gpu_amd import EVAL_BATCHED, EVAL_CACHED, EVAL_PLAIN
from symbolic_regression_gpu_amd.api import (CACHE_GENERATIONAL_REUSED, CACHE_REFERENCE, GaParams,
                                             SelectParams, read_data_file, read_digest)
from symbolic_regression_gpu_amd.replay import CSV_HEADER, evolve, replay
from tests.helpers import GRAMMAR_BINARY, GRAMMAR_UNARY, check_population, mse_close, uniform_data

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def test_trace_written_by_the_ga_replays_to_its_fitness(ctx, tmp_path):
    nv, N = 3, 6000
    X = uniform_data(nv, N, 12)
    X[nv] = X[0] * X[1] - np.cos(X[2])
    ctx.set_data(X)
    params = GaParams.make(num_vars=nv, max_tokens=80, max_depth=6, const_min=-2.0, const_max=2.0,
                           unary_ops=GRAMMAR_UNARY, binary_ops=GRAMMAR_BINARY)
    sel = SelectParams(7, 0.1, 0.05)
    trace = str(tmp_path / "trace")
    csv = str(tmp_path / "evolve.csv")
    rows = evolve(ctx, params, sel, 4, 9, 150, EVAL_BATCHED, csv_path=csv, trace_dir=trace, data_cols=X,
                  log=lambda *a: None)
    lines = open(csv).read().splitlines()
    assert lines[0] == CSV_HEADER and len(lines) == 5
    # the GA's own fitness per generation, re-evaluated from the trace files
    want = []
    for g in range(4):
        d = read_digest(os.path.join(trace, f"gen_{g}.txt"))
        assert d.shared and d.population.num_exprs == 150
        m, _, _ = ctx.eval_mse(d.population, EVAL_BATCHED)
        want.append(m)
        assert rows[g]["BestMSE"] == float(np.min(m[np.isfinite(m)]))
    cols, n = read_data_file(os.path.join(trace, "shared_data.txt"), nv, N)
    assert n == (nv + 1) * N and np.array_equal(cols.astype(np.float32), X)
    logs = []
    got_b = replay(trace, 0, 3, EVAL_BATCHED, ctx, log=logs.append)
    assert logs[0] == "Running Evolution Benchmark: Gen 0 to 3" and logs[-1].startswith("### Total Evolution Time:")
    assert all(r["summary"].startswith(f"Gen {r['gen']}: ") for r in got_b)
    try:
        ctx.cache_policy(CACHE_GENERATIONAL_REUSED, 0, 4)
        got_c = replay(trace, 0, 3, EVAL_CACHED, ctx, csv_path=str(tmp_path / "replay.csv"), log=lambda *a: None)
        ctx.cache_policy(CACHE_REFERENCE, 0, 4)
        ctx.cache_clear()
        got_r = replay(trace, 0, 3, EVAL_CACHED, ctx, log=lambda *a: None)
    finally:
        ctx.cache_policy(CACHE_REFERENCE, 0, 4)
    got_p = replay(trace, 0, 3, EVAL_PLAIN, ctx, log=lambda *a: None)
    for g in range(4):
        assert np.array_equal(got_b[g]["mse"], want[g], equal_nan=True), g
        assert np.array_equal(got_c[g]["mse"], want[g], equal_nan=True), g
        assert np.array_equal(got_r[g]["mse"], want[g], equal_nan=True), g
        assert mse_close(got_p[g]["mse"], want[g], 1e-5).all(), g
    assert len(open(tmp_path / "replay.csv").read().splitlines()) == 5
    # a missing generation is skipped, as the reference does
    assert [r["gen"] for r in replay(trace, 2, 5, EVAL_BATCHED, ctx, log=lambda *a: None)] == [2, 3]


def test_reference_example_digest_replays_to_oracle(ctx, tmp_path):
    """data/examples/sample_input.txt: two expressions over two data files."""
    src = os.path.join(HERE, "golden", "formats")
    (tmp_path / "data" / "examples").mkdir(parents=True)
    for f in ("sample_1", "sample_2"):
        shutil.copy(os.path.join(src, f), tmp_path / "data" / "examples" / f)
    trace = tmp_path / "trace"
    trace.mkdir()
    shutil.copy(os.path.join(src, "sample_input.txt"), trace / "gen_0.txt")
    out = replay(str(trace), 0, 0, EVAL_BATCHED, ctx, data_root=str(tmp_path), log=lambda *a: None)
    assert len(out) == 1
    d = read_digest(str(trace / "gen_0.txt"))
    for e, name in enumerate(("sample_1", "sample_2")):
        nv = int(d.num_vars[e])
        cols, _ = read_data_file(str(tmp_path / "data" / "examples" / name), nv, 1000)
        X = cols.astype(np.float32)
        pop = O.Population([d.population.expr(e)])
        check_population(pop, X, nv, 1000, out[0]["mse"][e:e + 1], out[0]["valid"][e:e + 1])
