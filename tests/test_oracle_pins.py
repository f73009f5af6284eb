"""Independent pins of the oracle functions the GPU's integer results are held
to, where the reference code cannot be compiled here:

* the cross-generation cache (cl_gpu_eval.cu:294-329 marking, :384-397 slot
  assignment, :433-459 skeletons, :165-186 skeleton interpretation) against
  brute-force models written below (every subtree compared with every cached
  body, no hashing);
* the selection loop (cl_sr_system.cpp:285-330, inside the reference's main())
  with scripted random draws: the oracle's choices are recomputed from the
  draws it made and its children are replayed through the mutate / crossover
  functions that the compiled reference pins (tests/test_oracle_ref.py).

CPU only."""
from __future__ import annotations

import math
import struct

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import (ALL_BINARY, ALL_UNARY, GRAMMAR_BINARY, GRAMMAR_UNARY, collision_population,
                           planted_population, random_population)


# ---- brute-force models ------------------------------------------------------
def _arity(t):
    """The interpreter's arity (cl_cpu_eval.cpp:82): leaves <= 0, 1..9 binary."""
    return 0 if t <= 0 else (2 if t < 10 else 1)


def _sizes(tok):
    """Subtree size at every position by recursive descent from the left."""
    tok = [int(t) for t in tok]
    size = [0] * len(tok)

    def walk(i):
        n, j = 1, i + 1
        for _ in range(_arity(tok[i])):
            c = walk(j)
            n += c
            j += c
        size[i] = n
        return n

    i = 0
    while i < len(tok):
        i += walk(i)
    return size


def _body(tok, val, p, n):
    return tuple((int(tok[k]), struct.pack("<d", float(val[k]))) for k in range(p, p + n))


def _exprs(pop):
    return [pop.expr(e) for e in range(pop.E)]


def model_cache_update(det, pop, bodies, capacity):
    """Slots in detected-subtree order, a body already cached gets none, none
    beyond capacity.  bodies: list of cached bodies (slot = index), updated."""
    new_sub, new_slot = [], []
    for i in range(len(det["sub_size"])):
        t, v = pop.expr(int(det["sub_expr"][i]))
        b = _body(t, v, int(det["sub_start"][i]), int(det["sub_size"][i]))
        if b in bodies or len(bodies) >= capacity:
            continue
        new_sub.append(i)
        new_slot.append(len(bodies))
        bodies.append(b)
    return new_sub, new_slot


def model_mark(pop, bodies, min_size):
    """Outermost-first marking: the root of every cached subtree of size >=
    min_size that no marked subtree contains gets its slot, its interior -2."""
    slot_of = {b: s for s, b in enumerate(bodies)}
    out = []
    for t, v in _exprs(pop):
        sz = _sizes(t)
        h = [-1] * len(t)
        p = 0
        while p < len(t):
            s = slot_of.get(_body(t, v, p, sz[p])) if sz[p] >= min_size else None
            if s is None:
                p += 1
                continue
            h[p] = s
            for k in range(p + 1, p + sz[p]):
                h[k] = -2
            p += sz[p]
        out.extend(h)
    return np.array(out, np.int32)


def _populations():
    yield "planted", planted_population(60, 5, nv=3, pool_size=8), 3, 5
    yield "planted-minfreq2", planted_population(40, 6, nv=4, pool_size=6, unary=GRAMMAR_UNARY,
                                                  binary=GRAMMAR_BINARY), 2, 2
    yield "random", random_population(50, 7, nv=2, max_tokens=40, max_depth=6, unary=GRAMMAR_UNARY,
                                      binary=GRAMMAR_BINARY), 3, 2
    yield "collision", collision_population(), 3, 2


@pytest.mark.parametrize("case", list(range(4)))
def test_cache_update_and_marking_equal_brute_force(case):
    name, pop, ms, mf = list(_populations())[case]
    ck = O.CacheKeys(10_000)
    bodies = []
    det = O.detect(pop, ms, mf)
    new_sub, new_slot = ck.update(pop, ms, mf)
    m_sub, m_slot = model_cache_update(det, pop, bodies, 10_000)
    assert new_sub.tolist() == m_sub and new_slot.tolist() == m_slot, name
    assert len(ck.keys) == len(bodies)
    # a later population (the generation after) is marked against the cache
    later = planted_population(30, 40 + case, nv=3, pool_size=5)
    mixed = O.Population(_exprs(pop)[:15] + _exprs(later))
    for p in (pop, mixed):
        for min_size in (ms, ms + 2):
            hints = ck.mark(p, min_size)
            ref = model_mark(p, bodies, min_size)
            assert np.array_equal(hints, ref), (name, min_size)
            # the hint map's invariants, spelled out
            for e, (t, v) in enumerate(_exprs(p)):
                sz = _sizes(t)
                h = hints[p.off[e]:p.off[e] + p.len[e]]
                covered = np.zeros(len(t), bool)
                for q in range(len(t)):
                    if h[q] >= 0:
                        assert sz[q] >= min_size and bodies[h[q]] == _body(t, v, q, sz[q])
                        assert not covered[q:q + sz[q]].any()  # spans never overlap
                        assert (h[q + 1:q + sz[q]] == -2).all()
                        covered[q:q + sz[q]] = True
                for q in range(len(t)):
                    if not covered[q] and sz[q] >= min_size:
                        assert _body(t, v, q, sz[q]) not in bodies  # nothing eligible left unmarked
                assert (h[~covered] == -1).all()
    ck.close()


def test_cache_update_respects_capacity_and_existing_bodies():
    pop = planted_population(60, 9, nv=3, pool_size=8)
    det = O.detect(pop, 3, 3)
    n = len(det["sub_size"])
    assert n >= 6
    cap = n // 2
    ck = O.CacheKeys(cap)
    bodies = []
    s1, l1 = ck.update(pop, 3, 3)
    m1 = model_cache_update(det, pop, bodies, cap)
    assert (s1.tolist(), l1.tolist()) == m1 and len(bodies) == cap
    # the same population again: every detected body is cached or the cache is full
    s2, _ = ck.update(pop, 3, 3)
    assert len(s2) == 0
    ck.close()
    # a cache holding some bodies already: those are skipped, the rest numbered on
    ck = O.CacheKeys(10_000)
    bodies = []
    first = O.Population(_exprs(pop)[:20])
    ck.update(first, 3, 3)
    model_cache_update(O.detect(first, 3, 3), first, bodies, 10_000)
    s3, l3 = ck.update(pop, 3, 3)
    assert (s3.tolist(), l3.tolist()) == model_cache_update(det, pop, bodies, 10_000)
    ck.close()


@pytest.mark.parametrize("case", list(range(4)))
def test_skeleton_expands_to_the_expression_and_evaluates_bit_exactly(case):
    name, pop, ms, mf = list(_populations())[case]
    nv, N = 4, 257
    X = np.random.default_rng(case).uniform(-2.5, 4.0, size=(nv + 1, N)).astype(np.float32)
    ck = O.CacheKeys(10_000)
    bodies = []
    det = O.detect(pop, ms, mf)
    ck.update(pop, ms, mf)
    model_cache_update(det, pop, bodies, 10_000)
    hints = ck.mark(pop, ms)
    sk_tok, sk_val, sk_off, sk_len = O.skeleton(pop, hints)
    # columns hold predict32 of each slot's body
    cache = np.zeros((max(len(bodies), 1), N), np.float32)
    for s, b in enumerate(bodies):
        bt = np.array([x[0] for x in b], np.int32)
        bv = np.array([struct.unpack("<d", x[1])[0] for x in b], np.float64)
        cache[s] = O.predict32(bt, bv, X, nv, N)
    n_ref = 0
    for e, (t, v) in enumerate(_exprs(pop)):
        st = sk_tok[sk_off[e]:sk_off[e] + sk_len[e]]
        sv = sk_val[sk_off[e]:sk_off[e] + sk_len[e]]
        # expanding every OP_REF by its slot's body gives back the stream
        # (values at fp32 precision, as the reference's skeleton packs them)
        et, ev = [], []
        for k in range(len(st)):
            if st[k] == -2:
                n_ref += 1
                for bt, bv in bodies[int(sv[k])]:
                    et.append(bt)
                    ev.append(float(np.float32(struct.unpack("<d", bv)[0])))
            else:
                et.append(int(st[k]))
                ev.append(float(sv[k]))
        assert et == [int(x) for x in t], (name, e)
        assert ev == [float(np.float32(x)) for x in v], (name, e)
        # and the skeleton over the columns evaluates to the expression, bit for bit
        a = O.predict32_cached(st, sv, X, nv, N, cache)
        b = O.predict32(t, v, X, nv, N)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (name, e)
    assert n_ref > 0, name
    ck.close()


# ---- selection loop ----------------------------------------------------------
class Script:
    """Per-child scripted draws: uniform / randint answer from a seeded stream
    and every call is logged as (kind, lo, hi, value)."""

    def __init__(self, seed, bias_first=None):
        self.seed = seed
        self.logs = {}
        self.bias_first = bias_first  # child -> forced first tournament draw

    def __call__(self, c):
        rng = np.random.default_rng(self.seed * 100_003 + c)
        log = self.logs.setdefault(c, [])

        def uniform(lo, hi):
            u = float(np.float32(lo + (hi - lo) * rng.random()))
            log.append(("u", lo, hi, u))
            return u

        def randint(lo, hi):
            if not log and self.bias_first is not None and c in self.bias_first and lo == 0:
                k = self.bias_first[c]
            else:
                k = int(rng.integers(lo, hi + 1)) if hi > lo else lo
            log.append(("i", lo, hi, k))
            return k

        return uniform, randint


def _replay(log):
    """RNG hooks that give back logged draws, checking each call's kind and range."""
    it = iter(log)

    def uniform(lo, hi):
        k, a, b, v = next(it)
        assert k == "u" and np.float32(a) == np.float32(lo) and np.float32(b) == np.float32(hi)
        return v

    def randint(lo, hi):
        k, a, b, v = next(it)
        assert k == "i" and (a, b) == (lo, hi)
        return v

    return uniform, randint, it


def _winner(draws, sfit):
    """Tournament as cl_sr_system.cpp:305-313 writes it: strict <, the earlier
    draw kept on ties (and a NaN drawn first is never displaced)."""
    best = draws[0]
    for c in draws[1:]:
        if sfit[c] < sfit[best]:
            best = c
    return best


@pytest.mark.parametrize("seed,tournament,p_mut,frac,pop_n", [(1, 7, 0.1, 0.05, 60), (2, 3, 0.2, 0.05, 40),
                                                              (3, 2, 0.5, 0.01, 30), (4, 1, 0.3, 0.2, 25)])
def test_selection_flow_from_scripted_draws(seed, tournament, p_mut, frac, pop_n):
    nv, max_tokens, max_depth = 3, 60, 6
    p = O.ga_params(num_vars=nv, max_tokens=max_tokens, max_depth=max_depth, const_min=-2.0, const_max=2.0,
                    unary_ops=ALL_UNARY, binary_ops=ALL_BINARY)
    cap = max_tokens + max_depth + 2
    base = random_population(pop_n, 50 + seed, nv=nv, max_tokens=max_tokens, max_depth=max_depth)
    tok = np.zeros(pop_n * cap, np.int32)
    val = np.zeros(pop_n * cap, np.float32)
    ln = np.zeros(pop_n, np.int32)
    for i in range(pop_n):
        t, v = base.expr(i)
        tok[i * cap:i * cap + len(t)] = t
        val[i * cap:i * cap + len(t)] = v.astype(np.float32)
        ln[i] = len(t)
    rng = np.random.default_rng(seed)
    # few distinct values, so tournaments often draw equal fitness (ties)
    fit = rng.integers(0, 5, pop_n).astype(np.float32) * np.float32(0.75)
    fit[rng.random(pop_n) < 0.2] = np.nan
    # order by an independent stable sort: ascending, NaN last, ties by index
    order = sorted(range(pop_n), key=lambda i: (math.isnan(fit[i]), 0.0 if math.isnan(fit[i]) else fit[i], i))
    assert O.fitness_order(fit).tolist() == order
    sfit = [float(fit[i]) for i in order]
    nan_rank = next(r for r in range(pop_n) if math.isnan(sfit[r]))
    script = Script(seed, bias_first={c: nan_rank for c in range(0, pop_n, 5)})
    rc, otok, oval, oln, p1, p2 = O.ga_next_generation_hooked(p, tournament, p_mut, frac, pop_n, cap, tok, val,
                                                              ln, fit, script)
    assert rc == 0
    elites = max(1, int(pop_n * frac))
    n_mut = n_cross = nan_first = ties = 0
    for c in range(pop_n):
        child = (otok[c * cap:c * cap + oln[c]].tolist(), oval[c * cap:c * cap + oln[c]].tolist())
        if c < elites:
            s = order[c]
            assert c not in script.logs  # elites draw nothing
            assert (p1[c], p2[c]) == (s, -2)
            assert child == (tok[s * cap:s * cap + ln[s]].tolist(), val[s * cap:s * cap + ln[s]].tolist())
            continue
        log = script.logs[c]
        draws = [e[3] for e in log[:tournament]]
        assert all(e[:3] == ("i", 0, pop_n - 1) for e in log[:tournament])
        w = _winner(draws, sfit)
        nan_first += math.isnan(sfit[draws[0]])
        ties += any(sfit[d] == sfit[w] and order[d] != order[w] for d in draws)
        if math.isnan(sfit[draws[0]]):
            assert w == draws[0]
        assert p1[c] == order[w]
        kind, lo, hi, u = log[tournament]
        assert kind == "u" and (lo, hi) == (0.0, 1.0)
        rest = log[tournament + 1:]
        a = (tok[p1[c] * cap:p1[c] * cap + ln[p1[c]]], val[p1[c] * cap:p1[c] * cap + ln[p1[c]]])
        if u < np.float32(p_mut):
            n_mut += 1
            assert p2[c] == -1
            uni, ri, it = _replay(rest)
            got = O.ga_op_with_hooks(1, p, a, a, uni, ri)
        else:
            n_cross += 1
            d2 = [e[3] for e in rest[:tournament]]
            assert all(e[:3] == ("i", 0, pop_n - 1) for e in rest[:tournament])
            assert p2[c] == order[_winner(d2, sfit)]
            b = (tok[p2[c] * cap:p2[c] * cap + ln[p2[c]]], val[p2[c] * cap:p2[c] * cap + ln[p2[c]]])
            uni, ri, it = _replay(rest[tournament:])
            got = O.ga_op_with_hooks(2, p, a, b, uni, ri)
        assert next(it, None) is None  # the child consumed every remaining draw
        assert (got[0].tolist(), got[1].tolist()) == child, c
    assert n_mut > 0 and n_cross > 0 and nan_first > 0 and (ties > 0 or tournament == 1)


def test_philox_binding_equals_hooked_loop_with_philox_draws():
    """sro_ga_next_generation (Philox streams, the GPU's) is the hooked loop
    with the stream (seed, child, gen, 1) bound to each child."""
    nv, cap, pop_n, seed, gen = 3, 68, 40, 77, 5
    p = O.ga_params(num_vars=nv, max_tokens=60, max_depth=6, unary_ops=GRAMMAR_UNARY, binary_ops=GRAMMAR_BINARY)
    rc, tok, val, ln = O.ga_init_population(p, pop_n, cap, 11)
    fit = np.random.default_rng(0).uniform(0, 5, pop_n).astype(np.float32)
    rc, t1, v1, l1 = O.ga_next_generation(p, 7, 0.1, 0.05, pop_n, cap, tok, val, ln, fit, seed, gen)
    streams = {}

    def child(c):
        st = O.PhiloxStream(seed, c, gen, 1)
        streams[c] = st
        return st.uniform, st.randint

    rc2, t2, v2, l2, _, _ = O.ga_next_generation_hooked(p, 7, 0.1, 0.05, pop_n, cap, tok, val, ln, fit, child)
    assert rc == rc2 == 0 and np.array_equal(l1, l2)
    for c in range(pop_n):
        n = l1[c]
        assert np.array_equal(t1[c * cap:c * cap + n], t2[c * cap:c * cap + n])
        assert np.array_equal(v1[c * cap:c * cap + n].view(np.uint32), v2[c * cap:c * cap + n].view(np.uint32))
