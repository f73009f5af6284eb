"""evogp population layout (defs.h:30-107) <-> prefix token streams, on the
oracle.

evogp -> prefix restates gpu_simple_jinha_with_evolve.cu:846-866: constants
and variables keep their value, a function node's value is its opcode, the
output flag (bit 7 of the type) is dropped, an individual whose size[0] is
outside 1..max_len is skipped.  The reverse direction derives node types from
token arity and subtree sizes; it is pinned here by hand-worked trees, by an
independent recursive parser and by the size invariants of prefix trees.
"""
import numpy as np

from oracle import oracle as O
from tests.helpers import GRAMMAR_BINARY, GRAMMAR_UNARY

VAR, CONST, UFUNC, BFUNC, OUT = 0, 1, 2, 3, 1 << 7


def _parse_sizes(tok):
    """Recursive-descent subtree sizes of one prefix stream."""
    sizes = [0] * len(tok)

    def node(i):
        k = tok[i]
        j = i + 1
        if k not in (0, -1):
            for _ in range(2 if 1 <= k <= 9 else 1):
                j = node(j)
        sizes[i] = j - i
        return j

    assert node(0) == len(tok)
    return sizes


def test_known_answer_both_directions():
    # add(x0, sin(2.0)): tokens [add, VAR, sin, CONST] with values [0, 0, 0, 2]
    tok = np.array([1, -1, 10, 0], np.int32)
    val = np.array([0, 0, 0, 2.0], np.float32)
    value, type_, size = O.prefix_to_evogp(tok, val, np.array([4], np.int32), 4, 6)
    assert list(value[:4]) == [1.0, 0.0, 10.0, 2.0]
    assert list(type_[:4]) == [BFUNC, VAR, UFUNC, CONST]
    assert list(size[:4]) == [4, 1, 2, 1]
    # and back, with the output flag set on the root (it is ignored)
    type_[0] |= OUT
    t2, v2, n2 = O.evogp_to_prefix(value, type_, size, 6, 8)
    assert n2[0] == 4 and list(t2[:4]) == [1, -1, 10, 0] and list(v2[:4]) == [0, 0, 0, 2.0]
    # x1 * (x0 - 0.5): variable index kept as the value of a VAR node
    value = np.array([3, 1, 2, 0, 0.5, 0, 0, 0], np.float32)
    type_ = np.array([BFUNC, VAR, BFUNC, VAR, CONST, 0, 0, 0], np.int16)
    size = np.array([5, 1, 3, 1, 1, 0, 0, 0], np.int16)
    t3, v3, n3 = O.evogp_to_prefix(value, type_, size, 8, 8)
    assert n3[0] == 5 and list(t3[:5]) == [3, -1, 2, -1, 0] and list(v3[:5]) == [0, 1, 0, 0, 0.5]


def test_skipped_individuals():
    max_len, cap = 8, 6
    value = np.zeros(4 * max_len, np.float32)
    type_ = np.full(4 * max_len, CONST, np.int16)
    size = np.zeros(4 * max_len, np.int16)
    size[[0, max_len, 2 * max_len, 3 * max_len]] = [0, -3, 7, 1]  # 7 > cap
    tok, val, ln = O.evogp_to_prefix(value, type_, size, max_len, cap)
    assert list(ln) == [0, 0, 0, 1]
    # reverse: too long, trailing tokens, missing operand, unknown token
    cases = [([0] * 9, 9), ([1, 0, 0, 0], 4), ([1, 0], 2), ([10, 28, 0], 3), ([-2], 1), ([10, 0], 2)]
    cap = 10
    tok = np.zeros(len(cases) * cap, np.int32)
    ln = np.zeros(len(cases), np.int32)
    for i, (t, n) in enumerate(cases):
        tok[i * cap:i * cap + len(t)] = t
        ln[i] = n
    value, type_, size = O.prefix_to_evogp(tok, np.zeros_like(tok, dtype=np.float32), ln, cap, 8)
    assert list(size[::8]) == [0, 0, 0, 0, 0, 2]


def test_round_trip_and_size_invariants_on_ga_populations():
    for seed, kw in [(3, dict(num_vars=3, max_tokens=255, max_depth=8)),
                     (9, dict(num_vars=5, max_tokens=64, max_depth=5, prob_leaf=0.3))]:
        p = O.ga_params(unary_ops=GRAMMAR_UNARY, binary_ops=GRAMMAR_BINARY, **kw)
        pop = 300
        cap = kw["max_tokens"] + kw["max_depth"] + 2
        rc, tok, val, ln = O.ga_init_population(p, pop, cap, seed)
        assert rc == 0
        max_len = int(ln.max())
        value, type_, size = O.prefix_to_evogp(tok, val, ln, cap, max_len)
        for i in range(pop):
            n = int(ln[i])
            t = tok[i * cap:i * cap + n]
            s = size[i * max_len:i * max_len + n]
            ty = type_[i * max_len:i * max_len + n]
            assert list(s) == _parse_sizes(list(t)), i
            for j in range(n):  # a binary node's size = 1 + both children
                if ty[j] == BFUNC:
                    assert s[j] == 1 + s[j + 1] + s[j + 1 + s[j + 1]]
                elif ty[j] == UFUNC:
                    assert s[j] == 1 + s[j + 1]
                else:
                    assert s[j] == 1
            assert np.array_equal(ty, np.where(t == 0, CONST, np.where(t == -1, VAR, np.where(
                (t >= 1) & (t <= 9), BFUNC, UFUNC))))
        t2, v2, n2 = O.evogp_to_prefix(value, type_, size, max_len, cap)
        assert np.array_equal(n2, ln)
        for i in range(pop):
            n = int(ln[i])
            a, b = slice(i * cap, i * cap + n), tok[i * cap:i * cap + n]
            assert np.array_equal(t2[a], b)
            leaf = (b == 0) | (b == -1)
            assert np.array_equal(v2[a][leaf].view(np.int32), val[a][leaf].view(np.int32))
            assert np.all(v2[a][~leaf] == 0)
