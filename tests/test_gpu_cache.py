"""Subtree-cached evaluation (mode CACHED) against the oracle.

* detection output (flat hashes, sizes, counts, hint map) and the cache's
  slot keys / per-position hit map equal the oracle's restatement of
  detect.cpp:120-387 and cl_gpu_eval.cu:294-329,384-397 exactly;
* cached MSEs equal uncached batched MSEs bit for bit (a cached column holds
  the value the full evaluation computes at that node), and the oracle's
  MSEs under the conditioning-aware bound of tests/helpers.py.
"""
import numpy as np
import pytest

from oracle import oracle as O
from symbolic_regression_gpu_amd import EVAL_BATCHED, EVAL_CACHED
from tests.helpers import (EPS32, GRAMMAR_BINARY, GRAMMAR_UNARY, check_population, collision_population,
                           mse_error_bound, random_population,
                           to_api, uniform_data)

pytestmark = pytest.mark.gpu


def mutants(base, n, seed, nv):
    """C3 shape: n copies of base, each with one subtree of depth 1..3
    replaced by a fresh random subtree of depth 1..3."""
    rng = np.random.default_rng(seed)
    bt, bv = base
    fresh = random_population(4 * n, seed + 1000, nv=nv, max_tokens=15, max_depth=3,
                              unary=GRAMMAR_UNARY, binary=GRAMMAR_BINARY)
    hs, sz = O.merkle(bt, bv)
    # depth of each subtree
    depth = np.zeros(len(bt), int)
    for i in range(len(bt) - 1, -1, -1):
        t = bt[i]
        d, pos = 1, i + 1
        for _ in range(0 if t <= 0 else (2 if t < 10 else 1)):
            d = max(d, 1 + depth[pos])
            pos += sz[pos]
        depth[i] = d
    cand = np.flatnonzero((depth >= 1) & (depth <= 3) & (np.arange(len(bt)) > 0))
    out = []
    for k in range(n):
        i = int(rng.choice(cand))
        ft, fv = fresh.expr(k)
        t = np.concatenate([bt[:i], ft, bt[i + sz[i]:]])
        v = np.concatenate([bv[:i], fv, bv[i + sz[i]:]])
        out.append((t, v))
    return out


def c3_population(seed=3):
    nv = 5
    base = random_population(1, seed, nv=nv, max_tokens=45, max_depth=7, unary=GRAMMAR_UNARY,
                             binary=GRAMMAR_BINARY, min_len=36)
    b = base.expr(0)
    return O.Population([b] + mutants(b, 100, seed, nv)), nv


def oracle_cache_replay(pops, capacity, min_size=3, min_freq=10):
    ck = O.CacheKeys(capacity)
    out = []
    for pop in pops:
        ck.update(pop, min_size, min_freq)
        out.append((ck.keys.copy(), ck.mark(pop, min_size)))
    ck.close()
    return out


def test_detect_matches_oracle(ctx):
    pop, nv = c3_population()
    X = uniform_data(nv, 1000, 1)
    ctx.set_data(X)
    for mn, mf in [(3, 10), (3, 2), (5, 20)]:
        got = ctx.detect(to_api(pop), mn, mf)
        ref = O.detect(pop, mn, mf)
        assert got["num_subs"] == ref["num_subs"]
        assert np.array_equal(got["flat"], ref["sub_flat"])
        assert np.array_equal(got["size"], ref["sub_size"])
        assert np.array_equal(got["count"], [len(o) for o in ref["occurrences"]])
        assert np.array_equal(got["hints"], ref["hints"])


@pytest.mark.parametrize("capacity", [2048, 7])
def test_cached_c3_matches_batched_and_oracle(ctx, capacity):
    pop, nv = c3_population()
    N = 200_001
    X = uniform_data(nv, N, 2)
    X[nv] = np.cos(X[0]) + X[1]
    ctx.set_data(X)
    ctx.cache_config(capacity, 3, 10)
    ctx.cache_clear()
    api = to_api(pop)
    m_b, v_b, _ = ctx.eval_mse(api, EVAL_BATCHED)
    m_c, v_c, st = ctx.eval_mse(api, EVAL_CACHED)
    assert np.array_equal(m_c, m_b, equal_nan=True)
    assert np.array_equal(v_c, v_b)
    check_population(pop, X, nv, N, m_c, v_c)
    keys, hints = oracle_cache_replay([pop], capacity)[0]
    assert np.array_equal(ctx.cache_keys(), keys)
    assert np.array_equal(ctx.cache_last_hints(), hints)
    assert st["num_cached"] == len(keys)
    assert st["executed_node_evals"] < st["logical_node_evals"]
    saved = 1 - st["executed_node_evals"] / st["logical_node_evals"]
    print(f"C3 capacity={capacity}: {st['num_subtrees']} detected, {st['num_cached']} cached, "
          f"coverage {st['coverage']:.3f}, node-evals saved {saved:.3f}")
    # second call: everything already cached, no new columns
    m_c2, _, st2 = ctx.eval_mse(api, EVAL_CACHED)
    assert np.array_equal(m_c2, m_c, equal_nan=True)
    assert st2["num_new_subtrees"] == 0


def test_cache_persists_across_generations(ctx):
    nv = 3
    X = uniform_data(nv, 30_000, 5)
    ctx.set_data(X)
    ctx.cache_config(64, 3, 10)
    ctx.cache_clear()
    base = random_population(12, 41, nv=nv, max_tokens=40, max_depth=6, unary=GRAMMAR_UNARY,
                             binary=GRAMMAR_BINARY, min_len=20)
    pops = []
    for g in range(3):
        exprs = []
        for k in range(60):
            b = base.expr((k + g) % 12)
            exprs += mutants(b, 1, 100 * g + k, nv)
        pops.append(O.Population(exprs))
    replay = oracle_cache_replay(pops, 64)
    for g, pop in enumerate(pops):
        api = to_api(pop)
        m_c, _, _ = ctx.eval_mse(api, EVAL_CACHED)
        m_b, _, _ = ctx.eval_mse(api, EVAL_BATCHED)
        assert np.array_equal(m_c, m_b, equal_nan=True), g
        keys, hints = replay[g]
        assert np.array_equal(ctx.cache_keys(), keys), g
        assert np.array_equal(ctx.cache_last_hints(), hints), g


# ---- generational policy ---------------------------------------------------
C4 = dict(num_vars=3, max_tokens=255, max_depth=8, const_min=-2.0, const_max=2.0,
          unary_ops=GRAMMAR_UNARY, binary_ops=GRAMMAR_BINARY)


@pytest.mark.parametrize("capacity,min_size,policy", [(0, 4, 1), (300, 2, 1), (0, 4, 2), (0, 3, 3), (200, 3, 3)])
def test_generational_cache_matches_batched_over_generations(ctx, capacity, min_size, policy):
    """Six GA generations evaluated with the generational cache equal the
    uncached batched MSEs bit for bit, including once the cache is full
    (capacity 300: the subtrees without a column are evaluated inline)."""
    from symbolic_regression_gpu_amd.api import CACHE_GENERATIONAL, CACHE_REFERENCE, GaParams, SelectParams
    N = 20_000
    X = uniform_data(3, N, 11)
    X[3] = X[0] * X[1] + np.sin(X[2])
    ctx.set_data(X)
    ctx.cache_policy(policy, capacity, min_size)
    try:
        seed = 21
        ctx.ga_init(GaParams.make(**C4), 400, seed)
        sel = SelectParams(7, 0.1, 0.05)
        saved = []
        for g in range(6):
            m_c, st = ctx.ga_evaluate(EVAL_CACHED)
            m_b, _ = ctx.ga_evaluate(EVAL_BATCHED)
            assert np.array_equal(m_c, m_b, equal_nan=True), g
            if capacity:
                assert st["num_cached"] <= capacity
            saved.append(1 - st["executed_node_evals"] / st["logical_node_evals"])
            if g == 0:
                # same population again: every column exists already
                m_c2, st2 = ctx.ga_evaluate(EVAL_CACHED)
                assert np.array_equal(m_c2, m_b, equal_nan=True)
                if capacity == 0 and policy == 1:
                    # only expressions below min_size (no column) run again
                    assert st2["num_new_subtrees"] == 0
                    ln = ctx.ga_population()[2]
                    assert st2["executed_node_evals"] == N * int(ln[ln < min_size].sum())
            ctx.ga_next(sel, m_b, seed, g)
        print(f"generational cache policy={policy} capacity={capacity}: node-evals saved per gen "
              + " ".join(f"{s:.2f}" for s in saved))
        if capacity == 0:
            assert min(saved[1:]) > 0.3
    finally:
        ctx.cache_policy(CACHE_REFERENCE, 0, 4)


def test_generational_cache_c3_and_oracle(ctx):
    from symbolic_regression_gpu_amd.api import CACHE_GENERATIONAL, CACHE_REFERENCE
    pop, nv = c3_population(5)
    N = 100_003
    X = uniform_data(nv, N, 3)
    X[nv] = np.cos(X[0]) + X[1]
    ctx.set_data(X)
    ctx.cache_policy(CACHE_GENERATIONAL, 0, 3)
    try:
        api = to_api(pop)
        m_b, v_b, _ = ctx.eval_mse(api, EVAL_BATCHED)
        m_c, v_c, st = ctx.eval_mse(api, EVAL_CACHED)
        assert np.array_equal(m_c, m_b, equal_nan=True)
        assert np.array_equal(v_c, v_b)
        check_population(pop, X, nv, N, m_c, v_c)
        # one column per distinct subtree body of size >= 3
        bodies = set()
        for e in range(pop.E):
            t, v = pop.expr(e)
            _, sz = O.merkle(t, v)
            for i in range(len(t)):
                if sz[i] >= 3:
                    bodies.add((tuple(t[i:i + sz[i]]), tuple(v[i:i + sz[i]])))
        assert st["num_cached"] == len(bodies)
        assert st["executed_node_evals"] < 0.7 * st["logical_node_evals"]
        # the same population again: every MSE is remembered, nothing runs
        m_c1, v_c1, st1 = ctx.eval_mse(api, EVAL_CACHED)
        assert np.array_equal(m_c1, m_b, equal_nan=True)
        assert np.array_equal(v_c1, v_b)
        assert st1["executed_node_evals"] == 0 and st1["num_new_subtrees"] == 0
        # one expression replaced by a copy of another: no new columns
        exprs = [pop.expr(e) for e in range(pop.E)]
        exprs[7] = exprs[0]
        pop2 = O.Population(exprs)
        m_c2, _, st2 = ctx.eval_mse(to_api(pop2), EVAL_CACHED)
        m_b2, _, _ = ctx.eval_mse(to_api(pop2), EVAL_BATCHED)
        assert np.array_equal(m_c2, m_b2, equal_nan=True)
        assert st2["num_new_subtrees"] == 0
        # new data empties the cache; results still match
        X2 = X.copy()
        X2[nv] = X[0] * 0.5
        ctx.set_data(X2)
        m_c3, _, st3 = ctx.eval_mse(api, EVAL_CACHED)
        m_b3, _, _ = ctx.eval_mse(api, EVAL_BATCHED)
        assert np.array_equal(m_c3, m_b3, equal_nan=True)
        assert st3["num_new_subtrees"] == st["num_new_subtrees"]
    finally:
        ctx.cache_policy(CACHE_REFERENCE, 0, 4)


def test_generational_cache_edge_cases(ctx):
    """Mirror-image subtrees that the reference's Merkle combine maps to one
    hash, duplicates within a population, a 1024-token chain (stores limited
    by the program length) and expressions below min_size."""
    from symbolic_regression_gpu_amd.api import CACHE_GENERATIONAL, CACHE_REFERENCE
    SUB, MUL, NEG, ABS, SIN = 2, 3, 25, 24, 10
    V = lambda k: (-1, float(k))
    def expr(*items):
        t = [i[0] if isinstance(i, tuple) else i for i in items]
        v = [i[1] if isinstance(i, tuple) else 0.0 for i in items]
        return np.array(t, np.int32), np.array(v, np.float64)
    a = expr(SUB, V(2), MUL, V(0), V(2))
    b = expr(SUB, V(0), MUL, V(2), V(0))
    ha, _ = O.merkle(*a)
    hb, _ = O.merkle(*b)
    assert ha[0] == hb[0]  # the reference's combine collides here
    chain = expr(*([NEG, ABS, SIN] * 341), V(1))
    assert len(chain[0]) == 1024
    exprs = [a, b, a, expr(SIN, V(0)), chain, b, chain]
    pop = O.Population(exprs)
    nv, N = 3, 5000
    X = uniform_data(nv, N, 9)
    ctx.set_data(X)
    ctx.cache_policy(CACHE_GENERATIONAL, 0, 4)
    try:
        api = to_api(pop)
        m_b, v_b, _ = ctx.eval_mse(api, EVAL_BATCHED)
        for _ in range(3):
            m_c, v_c, _ = ctx.eval_mse(api, EVAL_CACHED)
            assert np.array_equal(m_c, m_b, equal_nan=True)
            assert np.array_equal(v_c, v_b)
        assert m_b[0] != m_b[1]
    finally:
        ctx.cache_policy(CACHE_REFERENCE, 0, 4)


def test_reference_policy_flat_hash_collision(ctx):
    pop = collision_population()
    nv, N = 3, 4099
    X = uniform_data(nv, N, 4)
    ctx.set_data(X)
    ctx.cache_config(64, 3, 5)
    ctx.cache_clear()
    api = to_api(pop)
    m_b, v_b, _ = ctx.eval_mse(api, EVAL_BATCHED)
    for _ in range(2):
        m_c, v_c, st = ctx.eval_mse(api, EVAL_CACHED)
        assert np.array_equal(m_c, m_b, equal_nan=True)
        assert np.array_equal(v_c, v_b)
    keys, hints = oracle_cache_replay([pop], 64, 3, 5)[0]
    assert np.array_equal(ctx.cache_keys(), keys) and len(keys) == 4 and len(set(keys.tolist())) == 3
    assert np.array_equal(ctx.cache_last_hints(), hints)
    assert m_b[0] != m_b[1]
    ctx.cache_config(2048, 3, 10)
    ctx.cache_clear()


def test_generational_cache_vs_fp64_subtree_oracle(ctx):
    """The cached path over GA generations against the stateful FP64 CPU
    subtree evaluator (cpu_subtree.cpp:445-501, its own column cache carried
    across the same generations).  Every expression whose relative MSE error
    bound (first-order running bound of an fp32 evaluation, 2 ulp per
    transcendental) is <= 1e-6 must lie within 1e-5 of that FP64 MSE; every
    other one with a finite bound within twice its bound."""
    from symbolic_regression_gpu_amd.api import CACHE_GENERATIONAL, CACHE_REFERENCE, GaParams, SelectParams
    nv, N = 3, 4096
    X = uniform_data(nv, N, 12)
    X[nv] = np.sin(X[0]) * X[1] + X[2]
    X64 = X.astype(np.float64)
    ctx.set_data(X)
    ctx.cache_policy(CACHE_GENERATIONAL, 0, 3)
    st64 = O.SubtreeState()
    try:
        seed = 31
        ctx.ga_init(GaParams.make(**C4), 300, seed)
        sel = SelectParams(7, 0.1, 0.05)
        checked = 0
        for g in range(4):
            m_c, st = ctx.ga_evaluate(EVAL_CACHED)
            tok, val, ln, cap = ctx.ga_population()
            pop = O.Population([(tok[i * cap:i * cap + ln[i]], val[i * cap:i * cap + ln[i]].astype(np.float64))
                                for i in range(len(ln))])
            pred, new, _ = st64.predict64(pop, X64, nv, N)
            assert new > 0 or g > 0
            y = X64[nv]
            for e in range(pop.E):
                t, v = pop.expr(e)
                c32 = O.predict32(t, v, X, nv, N)
                v64, b64 = O.predict64_bound(t, v, X, nv, N)
                rb = mse_error_bound(v64, b64, c32, y)
                fin = np.isfinite(c32)
                if not np.isfinite(rb) or not np.isfinite(m_c[e]):
                    continue
                m64 = float(np.mean((pred[e][fin] - y[fin]) ** 2))
                tol = 1e-5 if rb <= 1e-6 else 2 * rb + 4 * EPS32
                checked += rb <= 1e-6
                assert abs(float(m_c[e]) - m64) <= tol * m64, (g, e, m_c[e], m64, rb)
            ctx.ga_next(sel, m_c, seed, g)
        assert checked > 0.15 * 4 * 300
    finally:
        st64.close()
        ctx.cache_policy(CACHE_REFERENCE, 0, 4)


@pytest.mark.parametrize("nv", [3, 4])
def test_shared_policy_matches_batched_per_variable_class(ctx, nv):
    """Cache skeletons run the 3-register kernel for nv <= 3 and the
    4-register one for nv = 4; both equal the batched MSEs bit for bit."""
    from symbolic_regression_gpu_amd.api import CACHE_GENERATIONAL_SHARED, CACHE_REFERENCE
    base = random_population(1, 40 + nv, nv=nv, max_tokens=45, max_depth=7, unary=GRAMMAR_UNARY,
                             binary=GRAMMAR_BINARY, min_len=30)
    b = base.expr(0)
    pop = O.Population([b] + mutants(b, 150, 40 + nv, nv))
    N = 30_011
    X = uniform_data(nv, N, 6)
    ctx.set_data(X)
    ctx.cache_policy(CACHE_GENERATIONAL_SHARED, 0, 3)
    try:
        api = to_api(pop)
        m_b, v_b, _ = ctx.eval_mse(api, EVAL_BATCHED)
        for _ in range(2):
            m_c, v_c, st = ctx.eval_mse(api, EVAL_CACHED)
            assert np.array_equal(m_c, m_b, equal_nan=True)
            assert np.array_equal(v_c, v_b)
        assert st["executed_node_evals"] < st["logical_node_evals"]
        check_population(pop, X, nv, N, m_c, v_c)
    finally:
        ctx.cache_policy(CACHE_REFERENCE, 0, 4)
