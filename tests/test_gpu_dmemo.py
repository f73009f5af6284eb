"""The generational subtree cache planned on the device (SR_CACHE_DEVICE):
over GA generations its fitness vectors equal the uncached batched run's bit
for bit (columns hold exactly what the interpreter computes for a subtree),
with one captured graph per generation or without graphs, under eviction
pressure (a few slots), for several minimum subtree sizes and variable
counts; and it executes only a fraction of the logical tokens."""
import numpy as np
import pytest

from symbolic_regression_gpu_amd import EVAL_BATCHED, EVAL_CACHED
from symbolic_regression_gpu_amd.api import CACHE_DEVICE, CACHE_REFERENCE, GaParams, SelectParams, Tuning
from tests.helpers import GRAMMAR_BINARY, GRAMMAR_UNARY, uniform_data

pytestmark = pytest.mark.gpu


def _params(nv, max_tokens=120, max_depth=8):
    return GaParams.make(num_vars=nv, max_tokens=max_tokens, max_depth=max_depth, const_min=-2.0,
                         const_max=2.0, unary_ops=GRAMMAR_UNARY, binary_ops=GRAMMAR_BINARY)


def _runs(ctx, nv, pop, gens, seed, capacity, min_size, use_graph, N=20_011, max_tokens=120):
    X = uniform_data(nv, N, seed)
    X[nv] = np.sin(X[0]) * X[min(1, nv - 1)] + 1.5
    ctx.set_data(X)
    ctx.ga_init(_params(nv, max_tokens), pop, seed)
    start = ctx.ga_population()
    sel = SelectParams(7, 0.1, 0.05)
    ctx.cache_policy(CACHE_REFERENCE, 0, 3)
    hist_b, _, st_b = ctx.ga_run(sel, EVAL_BATCHED, gens, 0, seed, True)
    end_b = ctx.ga_population()
    ctx.ga_set_population(*start)
    ctx.cache_policy(CACHE_DEVICE, capacity, min_size)
    try:
        hist_c, _, st_c = ctx.ga_run(sel, EVAL_CACHED, gens, 0, seed, use_graph)
        end_c = ctx.ga_population()
    finally:
        ctx.cache_policy(CACHE_REFERENCE, 0, 3)
    return hist_b, st_b, end_b, hist_c, st_c, end_c


def _same(h1, h2):
    return np.array_equal(np.asarray(h1).view(np.uint32), np.asarray(h2).view(np.uint32))


@pytest.mark.parametrize("use_graph,cache_kernel", [(True, 0), (False, 0), (True, 1), (False, 1)])
def test_device_cache_equals_batched_over_generations(ctx, use_graph, cache_kernel):
    saved = ctx.tuning
    try:
        ctx.tuning = Tuning(saved.dpt, saved.chunk, saved.n_queues, saved.blocks_per_cu, saved.cache_dpt,
                            saved.kernel, cache_kernel)
        hb, sb, eb, hc, sc, ec = _runs(ctx, 3, 400, 8, 11, 0, 3, use_graph)
    finally:
        ctx.tuning = saved
    assert _same(hb, hc)
    assert np.array_equal(eb[2], ec[2])
    ex = sc["executed_node_evals"] / sc["logical_node_evals"]
    print(f"device cache: executed / logical {ex:.3f}, column reads {sc['column_reads'] // 20_011}, "
          f"writes {sc['column_writes'] // 20_011}, live {sc['num_cached']}, "
          f"kernel ms {sc['interp_kernel_ms']:.2f}")
    assert 0 < ex < 0.6 and sc["column_reads"] > 0 and sc["column_writes"] > 0
    if not use_graph:
        assert sc["interp_kernel_ms"] > 0 and sc["interp_launches"] == 16


@pytest.mark.parametrize("capacity", [8, 64])
def test_device_cache_under_eviction_pressure(ctx, capacity):
    hb, _, _, hc, sc, _ = _runs(ctx, 3, 300, 6, 23, capacity, 3, True)
    assert _same(hb, hc)
    assert sc["num_cached"] <= capacity


@pytest.mark.parametrize("min_size,nv", [(2, 1), (4, 4), (6, 5), (3, 10)])
def test_device_cache_min_size_and_variable_classes(ctx, min_size, nv):
    hb, _, _, hc, sc, _ = _runs(ctx, nv, 250, 5, 31 + nv, 0, min_size, True)
    assert _same(hb, hc)
    assert sc["executed_node_evals"] < sc["logical_node_evals"]


def test_device_cache_evaluate_twice_and_duplicates(ctx):
    """Evaluating one population twice (columns of the first pass read by the
    second) and a population of many identical expressions (evaluated once,
    then settled by its remembered MSE)."""
    nv, N = 3, 9_001
    X = uniform_data(nv, N, 5)
    ctx.set_data(X)
    ctx.ga_init(_params(nv), 200, 5)
    tok, val, ln, cap = ctx.ga_population()
    # every fourth individual copies individual 0
    for i in range(0, 200, 4):
        tok[i * cap:(i + 1) * cap] = tok[:cap]
        val[i * cap:(i + 1) * cap] = val[:cap]
        ln[i] = ln[0]
    ctx.ga_set_population(tok, val, ln, cap)
    m_ref, _ = ctx.ga_evaluate(EVAL_BATCHED)
    ctx.cache_policy(CACHE_DEVICE, 0, 3)
    try:
        m1, st1 = ctx.ga_evaluate(EVAL_CACHED)
        m2, st2 = ctx.ga_evaluate(EVAL_CACHED)
    finally:
        ctx.cache_policy(CACHE_REFERENCE, 0, 3)
    assert _same(m1, m_ref) and _same(m2, m_ref)
    assert st2["executed_node_evals"] < st1["executed_node_evals"]


def test_skeleton_tuner_keeps_results_and_persists(ctx, tmp_path):
    """sr_autotune_cache refuses to run before a cached generation, then times
    the skeleton kernels on the last plan; the generations that follow keep
    the uncached run's fitness bit for bit, and the remembered tuning is
    written to and read back from a file."""
    from symbolic_regression_gpu_amd import SrError, api
    nv, pop, gens, seed = 3, 300, 6, 77
    X = uniform_data(nv, 20_011, seed)
    X[nv] = np.sin(X[0]) * X[1] + 1.5
    ctx.set_data(X)
    ctx.ga_init(_params(nv), pop, seed)
    start = ctx.ga_population()
    sel = SelectParams(7, 0.1, 0.05)
    saved = ctx.tuning
    try:
        ctx.cache_policy(CACHE_REFERENCE, 0, 3)
        hist_b, _, _ = ctx.ga_run(sel, EVAL_BATCHED, gens, 0, seed, True)
        ctx.ga_set_population(*start)
        ctx.cache_policy(CACHE_DEVICE, 0, 3)
        with pytest.raises(SrError):
            ctx.autotune_cache()
        h1, _, _ = ctx.ga_run(sel, EVAL_CACHED, 2, 0, seed, True)
        t, ms = ctx.autotune_cache()
        assert ms > 0 and t.cache_kernel in (0, 1) and 0 <= t.cache_dpt <= 4
        assert ctx.tuning.cache_kernel == t.cache_kernel and ctx.tuning.cache_dpt == t.cache_dpt
        h2, _, _ = ctx.ga_run(sel, EVAL_CACHED, gens - 2, 2, seed, True)
        assert _same(np.concatenate([h1, h2]), hist_b)
        path = tmp_path / "tune.txt"
        api.tune_save(str(path))
        rows = [line.split() for line in path.read_text().splitlines() if line and not line.startswith("#")]
        mine = [r for r in rows if r[0].startswith("gfx") and int(r[2]) == 4 and int(r[3]) == 20_011]
        assert mine and int(mine[0][9]) in (0, 1) and int(mine[0][10]) == t.cache_kernel
        assert api.tune_load(str(path)) == len(rows)
    finally:
        ctx.cache_policy(CACHE_REFERENCE, 0, 3)
        ctx.tuning = saved
