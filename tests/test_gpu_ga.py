"""GPU genetic algorithm vs the oracle's restatement (which is itself pinned
bit-for-bit to the reference's generate/mutate/crossover under mt19937, see
test_oracle_ref.py).  Offspring must match bit-exactly given the same seed,
generation index and fitness vector."""
import numpy as np
import pytest

from oracle import oracle as O
from symbolic_regression_gpu_amd import EVAL_BATCHED, EVAL_PLAIN
from symbolic_regression_gpu_amd.api import GaParams, SelectParams
from tests.helpers import GRAMMAR_BINARY, GRAMMAR_UNARY, uniform_data

pytestmark = pytest.mark.gpu

C4 = dict(num_vars=3, max_tokens=255, max_depth=8, const_min=-2.0, const_max=2.0,
          unary_ops=GRAMMAR_UNARY, binary_ops=GRAMMAR_BINARY)


def _params(**kw):
    return GaParams.make(**kw), O.ga_params(**kw)


def _same_pop(a, b, cap):
    at, av, al = a
    bt, bv, bl = b
    assert np.array_equal(al, bl)
    for i in range(len(al)):
        n = al[i]
        assert np.array_equal(at[i * cap:i * cap + n], bt[i * cap:i * cap + n]), i
        assert np.array_equal(av[i * cap:i * cap + n].view(np.int32),
                              bv[i * cap:i * cap + n].view(np.int32)), i


@pytest.mark.parametrize("kw", [C4, dict(num_vars=5, max_tokens=64, max_depth=5),
                                dict(num_vars=1, max_tokens=31, max_depth=6, prob_leaf=0.3)])
def test_init_population_bit_exact(ctx, kw):
    gp, op = _params(**kw)
    pop, seed = 700, 1234
    ctx.ga_init(gp, pop, seed)
    tok, val, ln, cap = ctx.ga_population()
    rc, otok, oval, oln = O.ga_init_population(op, pop, cap, seed)
    assert rc == 0
    _same_pop((tok, val, ln), (otok, oval, oln), cap)
    assert ln.max() <= kw["max_tokens"] + kw["max_depth"]
    for i in range(0, pop, 50):
        assert O.stack_depth(tok[i * cap:i * cap + ln[i]]) > 0


def test_next_generation_bit_exact_with_nan_and_ties(ctx):
    gp, op = _params(**C4)
    pop, seed = 1000, 99
    ctx.ga_init(gp, pop, seed)
    tok, val, ln, cap = ctx.ga_population()
    rng = np.random.default_rng(3)
    fit = rng.uniform(0, 10, size=pop).astype(np.float32)
    fit[rng.integers(0, pop, 40)] = np.nan
    fit[rng.integers(0, pop, 40)] = 1.0  # ties
    sel = SelectParams(7, 0.1, 0.05)
    for gen in (0, 5):
        ctx.ga_set_population(tok, val, ln, cap)
        ctx.ga_next(sel, fit, seed, gen)
        g = ctx.ga_population()
        rc, ot, ov, ol = O.ga_next_generation(op, 7, 0.1, 0.05, pop, cap, tok, val, ln, fit, seed, gen)
        assert rc == 0
        _same_pop(g[:3], (ot, ov, ol), cap)
    # elites are the 50 best in (fitness, index) order
    order = O.fitness_order(fit)
    for k in range(50):
        e = order[k]
        assert np.array_equal(g[0][k * cap:k * cap + g[2][k]], tok[e * cap:e * cap + ln[e]])


@pytest.mark.parametrize("use_graph", [True, False])
def test_evolution_loop_matches_oracle_replay(ctx, use_graph):
    # 4 generations on the GPU; replay them in the oracle with the GPU's own
    # fitness vectors; the final populations must be identical.
    gp, op = _params(**C4)
    pop, seed, gens = 256, 7, 4
    N = 4096
    X = uniform_data(3, N, 4)
    ctx.set_data(X)
    ctx.ga_init(gp, pop, seed)
    t0, v0, l0, cap = ctx.ga_population()
    sel = SelectParams(7, 0.1, 0.05)
    hist, ms, st = ctx.ga_run(sel, EVAL_BATCHED, gens, 0, seed, use_graph=use_graph)
    final = ctx.ga_population()
    tok, val, ln = t0, v0, l0
    for g in range(gens):
        rc, tok, val, ln = O.ga_next_generation(op, 7, 0.1, 0.05, pop, cap, tok, val, ln,
                                                hist[g], seed, g)
        assert rc == 0
    _same_pop(final[:3], (tok, val, ln), cap)
    assert st["logical_node_evals"] > 0 and ms > 0


def test_graph_and_stream_runs_agree(ctx):
    gp, _ = _params(**C4)
    X = uniform_data(3, 8192, 2)
    ctx.set_data(X)
    sel = SelectParams(7, 0.1, 0.05)
    out = []
    for use_graph in (True, False, True):
        ctx.ga_init(gp, 300, 42)
        hist, _, _ = ctx.ga_run(sel, EVAL_BATCHED, 3, 0, 42, use_graph=use_graph)
        out.append((hist, ctx.ga_population()))
    for h, p in out[1:]:
        assert np.array_equal(h, out[0][0], equal_nan=True)
        assert np.array_equal(p[2], out[0][1][2])


def test_graph_recaptured_after_a_buffer_grows():
    """A captured generation bakes in the context's evaluation buffers; when
    an evaluation in between grows any of them, the next graph run must
    capture again (and still match the stream run).  A context of its own, so
    the buffer sizes are this test's."""
    from symbolic_regression_gpu_amd import Context
    from tests.helpers import random_population, to_api
    ctx = Context(0)
    gp, _ = _params(**C4)
    X = uniform_data(3, 8192, 3)
    ctx.set_data(X)
    sel = SelectParams(7, 0.1, 0.05)
    ctx.ga_init(gp, 300, 9)
    start = ctx.ga_population()
    _, _, st1 = ctx.ga_run(sel, EVAL_BATCHED, 2, 0, 9, use_graph=True)
    ctx.ga_set_population(*start)
    h2, _, st2 = ctx.ga_run(sel, EVAL_BATCHED, 2, 0, 9, use_graph=True)
    assert st2["graph_captures"] == 0  # same buffers: replayed
    # 1.25x + 40 expressions: the count / partial buffers grow
    big = random_population(300 * 5 // 4 + 40, 3, nv=3, max_tokens=30, max_depth=5)
    ctx.eval_mse(to_api(big), EVAL_BATCHED)
    ctx.ga_set_population(*start)
    h3, _, st3 = ctx.ga_run(sel, EVAL_BATCHED, 2, 0, 9, use_graph=True)
    assert st3["graph_captures"] == 2
    ctx.ga_set_population(*start)
    h4, _, _ = ctx.ga_run(sel, EVAL_BATCHED, 2, 0, 9, use_graph=False)
    assert np.array_equal(h2, h3, equal_nan=True) and np.array_equal(h3, h4, equal_nan=True)
    assert st1["graph_captures"] == 2
    ctx.close()


def test_host_mode_evaluation_matches_device(ctx):
    gp, _ = _params(**C4)
    X = uniform_data(3, 5000, 8)
    ctx.set_data(X)
    ctx.ga_init(gp, 200, 5)
    m_dev, _ = ctx.ga_evaluate(EVAL_BATCHED)
    m_plain, _ = ctx.ga_evaluate(EVAL_PLAIN)
    fin = np.isfinite(m_dev)
    assert np.array_equal(fin, np.isfinite(m_plain))
    assert np.allclose(m_dev[fin], m_plain[fin], rtol=1e-5)


# Known answers of the reference's evolution_tests.cpp (mutate at fixed node,
# crossover at fixed nodes), expressed on prefix tokens.
ADD, SUB, MUL, POW, NEG = 1, 2, 3, 5, 25


def _tree(items):
    t = [x if isinstance(x, int) and x > 0 else 0 for x in items]
    v = [0.0 if (isinstance(x, int) and x > 0) else float(x[0]) for x in items]
    return t, v


def C_(x):
    return (x,)


KATS = [
    # (left, left node, right, right node, expected value)   evolution_tests.cpp
    ([ADD, MUL, C_(2), C_(3), C_(4)], 0, [SUB, C_(8), C_(5)], 0, 3.0),                    # :436-477
    ([ADD, MUL, SUB, C_(5), C_(2), C_(3), C_(4)], 2, [ADD, C_(1), C_(1)], 0, 10.0),       # :482-525
    ([ADD, MUL, C_(1), C_(2), C_(3)], 0, [SUB, POW, C_(2), C_(3), C_(5)], 0, 3.0),        # :620-667
    ([ADD, MUL, C_(1), C_(2), SUB, C_(4), C_(3)], 4,
     [ADD, POW, C_(2), C_(3), MUL, C_(5), C_(2)], 4, 12.0),                               # :829-876
    ([ADD, MUL, ADD, C_(1), C_(2), SUB, C_(5), C_(3), NEG, C_(4)], 1,
     [SUB, POW, ADD, C_(2), C_(1), C_(3), MUL, C_(6), C_(2)], 1, 23.0),                   # :672-751
    ([MUL, ADD, C_(1), C_(2), SUB, C_(7), C_(4)], 0,
     [ADD, POW, C_(2), C_(3), MUL, C_(5), C_(2)], 0, 18.0),                               # :756-824
]


def test_splice_known_answers(ctx):
    cap = 32
    n = len(KATS)
    lt = np.zeros(n * cap, np.int32)
    lv = np.zeros(n * cap, np.float32)
    rt = np.zeros(n * cap, np.int32)
    rv = np.zeros(n * cap, np.float32)
    ll = np.zeros(n, np.int32)
    rl = np.zeros(n, np.int32)
    li = np.zeros(n, np.int32)
    ri = np.zeros(n, np.int32)
    for k, (L, a, R, b, _) in enumerate(KATS):
        t, v = _tree(L)
        lt[k * cap:k * cap + len(t)] = t
        lv[k * cap:k * cap + len(t)] = v
        ll[k] = len(t)
        t, v = _tree(R)
        rt[k * cap:k * cap + len(t)] = t
        rv[k * cap:k * cap + len(t)] = v
        rl[k] = len(t)
        li[k], ri[k] = a, b
    sel = np.arange(n, dtype=np.int32)
    ot, ov, ol = ctx.tree_splice(cap, (lt, lv, ll), sel, li, (rt, rv, rl), sel, ri)
    X = np.zeros((1, 1), np.float32)
    for k, (*_, want) in enumerate(KATS):
        t = ot[k * cap:k * cap + ol[k]]
        v = ov[k * cap:k * cap + ol[k]].astype(np.float64)
        assert O.predict32(t, v, X, 0, 1)[0] == np.float32(want), k
    # out-of-range node index -> copy of the left tree
    ot2, _, ol2 = ctx.tree_splice(cap, (lt, lv, ll), sel[:1], np.array([99], np.int32),
                                  (rt, rv, rl), sel[:1], np.array([0], np.int32))
    assert ol2[0] == ll[0] and np.array_equal(ot2[:ll[0]], lt[:ll[0]])
