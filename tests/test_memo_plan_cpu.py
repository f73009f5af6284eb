"""Generational-cache planner (host logic, no device) checked symbolically.

A column is modelled by the token body of the subtree it holds: replaying a
plan -- OP_REF (-2) splices the body of a column written by an EARLIER launch
or call, OP_STORE (-3) records the body of its operand for later launches --
must rebuild every launched expression exactly, and every remembered MSE must
belong to a slot whose column holds that very expression.  A plan that reads
a column in the launch that writes it, reads an evicted or stale slot, or
points an expression at the wrong program fails here.  Populations come from
the oracle's GA (bit-exact to the device GA), over several generations.
"""
import numpy as np
import pytest

from oracle import oracle as O
from symbolic_regression_gpu_amd.api import (CACHE_GENERATIONAL, CACHE_GENERATIONAL_REUSED,
                                             CACHE_GENERATIONAL_SHARED, MemoPlanner,
                                             Population)
from tests.helpers import GRAMMAR_BINARY, GRAMMAR_UNARY

OP_REF, OP_STORE = -2, -3
C4 = dict(num_vars=3, max_tokens=255, max_depth=8, const_min=-2.0, const_max=2.0,
          unary_ops=GRAMMAR_UNARY, binary_ops=GRAMMAR_BINARY)


def arity(t):
    return 0 if t <= 0 else (2 if t < 10 else 1)


def replay_program(tok, val, cols, stored):
    """Rebuild a skeleton program into (tokens, values) of the expression it
    evaluates (right-to-left stack pass); OP_STORE bodies go to `stored`
    (slot -> body)."""
    stack = []
    for i in range(len(tok) - 1, -1, -1):
        t, v = int(tok[i]), float(val[i])
        if t == OP_REF:
            stack.append(tuple(tuple(x) for x in cols[int(v)]))  # KeyError: never written
        elif t == OP_STORE:
            stored[int(v)] = stack[-1]
        else:
            bt, bv = [t], [v]
            for _ in range(arity(t)):
                ct, cv = stack.pop()
                bt += ct
                bv += cv
            stack.append((tuple(bt), tuple(bv)))
    assert len(stack) == 1
    return stack[0]


def check_plan(pop, plan, cols, memo, capacity):
    """Replay launches in order; returns per-program bodies.  cols / memo are
    updated like the device columns / remembered MSEs."""
    lb = plan["launch_begin"]
    bodies = []
    for k in range(len(lb) - 1):
        stored = {}
        for i in range(lb[k], lb[k + 1]):
            a, n = plan["offsets"][i], plan["lengths"][i]
            assert 1 <= n <= 1024
            t, v = plan["tokens"][a:a + n], plan["values"][a:a + n]
            bodies.append(replay_program(t, v, cols, stored))
        for s in stored:
            assert 0 <= s < capacity
        cols.update({s: (list(b[0]), list(b[1])) for s, b in stored.items()})
        for s in stored:  # a rewritten slot forgets its old MSE
            memo.pop(s, None)
    for i, body in enumerate(bodies):
        s = plan["root_slot"][i]
        if s >= 0:
            assert tuple(cols[s][0]) == body[0] and tuple(cols[s][1]) == body[1]
    for e in range(pop.num_exprs):
        t, v = pop.expr(e)
        want = (tuple(int(x) for x in t), tuple(float(x) for x in v))
        src = plan["src"][e]
        if src >= 0:
            assert bodies[src] == want, e
        else:
            s = -2 - src
            assert (tuple(cols[s][0]), tuple(cols[s][1])) == want, e
            assert memo[s] == fake_mse(want)
    return bodies


def fake_mse(body):
    return (float(np.float32(hash(body) % 100003)), len(body[0]))


def ga_populations(gens, pop, seed, cap=263):
    op = O.ga_params(**C4)
    rc, tok, val, ln = O.ga_init_population(op, pop, cap, seed)
    assert rc == 0
    rng = np.random.default_rng(seed)
    for g in range(gens):
        yield Population.from_strided(tok, val, ln, cap)
        fit = rng.uniform(0, 10, pop).astype(np.float32)
        rc, tok, val, ln = O.ga_next_generation(op, 7, 0.1, 0.05, pop, cap, tok, val, ln, fit, seed, g)
        assert rc == 0


def run_generations(planner, pops, capacity):
    cols, memo = {}, {}
    stats = []
    for pop in pops:
        plan = planner.plan(pop)
        bodies = check_plan(pop, plan, cols, memo, capacity)
        mse = np.array([fake_mse(b)[0] for b in bodies], np.float32)
        cnt = np.array([fake_mse(b)[1] for b in bodies], np.int64)
        planner.commit(mse, cnt)
        for i, b in enumerate(bodies):
            s = plan["root_slot"][i]
            if s >= 0:
                memo[s] = fake_mse(b)
        for e in range(pop.num_exprs):
            if plan["src"][e] < 0:
                s = -2 - plan["src"][e]
                m, k = planner.remembered(s)
                assert (m, k) == memo[s]
        stats.append((len(plan["tokens"]), pop.total_tokens, len(plan["launch_begin"]) - 1))
    return stats


@pytest.mark.parametrize("policy", [CACHE_GENERATIONAL, CACHE_GENERATIONAL_REUSED, CACHE_GENERATIONAL_SHARED])
@pytest.mark.parametrize("capacity,min_size", [(1 << 16, 4), (1 << 16, 2), (150, 3)])
def test_plans_rebuild_every_expression(policy, capacity, min_size):
    planner = MemoPlanner(capacity, min_size, policy)
    stats = run_generations(planner, ga_populations(5, 300, 17), capacity)
    if capacity > 1000:
        # later generations execute a fraction of their tokens
        assert all(ex < 0.8 * tot for ex, tot, _ in stats[2:])


def test_repeated_population_runs_nothing():
    planner = MemoPlanner(1 << 14, 4)
    pop = next(ga_populations(1, 200, 5))
    cols, memo = {}, {}
    for k in range(3):
        plan = planner.plan(pop)
        bodies = check_plan(pop, plan, cols, memo, 1 << 14)
        mse = np.array([fake_mse(b)[0] for b in bodies], np.float32)
        cnt = np.array([fake_mse(b)[1] for b in bodies], np.int64)
        planner.commit(mse, cnt)
        for i, b in enumerate(bodies):
            if plan["root_slot"][i] >= 0:
                memo[plan["root_slot"][i]] = fake_mse(b)
        if k > 0:
            # only expressions below min_size run again
            small = [e for e in range(pop.num_exprs) if pop.lengths[e] < 4]
            assert len(plan["offsets"]) == len(small)


def test_mirrored_trees_and_long_chain():
    """The reference's Merkle combine maps the mirrored pair to one hash; the
    planner must keep them apart.  A 1024-token chain cannot take a store
    everywhere (programs stay within 1024 tokens)."""
    SUB, MUL, NEG, ABS, SIN = 2, 3, 25, 24, 10
    a = ([SUB, -1, MUL, -1, -1], [0, 2, 0, 0, 2])
    b = ([SUB, -1, MUL, -1, -1], [0, 0, 0, 2, 0])
    ha, _ = O.merkle(np.array(a[0], np.int32), np.array(a[1], np.float64))
    hb, _ = O.merkle(np.array(b[0], np.int32), np.array(b[1], np.float64))
    assert ha[0] == hb[0]
    chain = ([NEG, ABS, SIN] * 341 + [-1], [0.0] * 1023 + [1.0])
    pop = Population.from_exprs([a, b, a, chain, b, chain, ([SIN, -1], [0, 0])])
    planner = MemoPlanner(4096, 4)
    run_generations(planner, [pop, pop, pop], 4096)


def test_full_cache_and_eviction():
    """Capacity 8: slots are recycled as populations change; every plan still
    rebuilds its expressions."""
    planner = MemoPlanner(8, 3)
    run_generations(planner, ga_populations(6, 60, 23), 8)


def test_rejects_internal_tokens():
    from symbolic_regression_gpu_amd.api import SrError
    planner = MemoPlanner(16, 3)
    for bad in (([-2], [0.0]), ([-3, 10, -1], [0.0, 0.0, 0.0]), ([1, -1], [0.0, 0.0])):
        with pytest.raises(SrError):
            planner.plan(Population.from_exprs([bad]))
