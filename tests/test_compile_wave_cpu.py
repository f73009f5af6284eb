"""The device-planned cache compiles its skeletons with a wave-parallel
compiler (csrc/sr_compile_wave.h: prefix scans for sizes and instruction
counts, rounds for the Sethi-Ullman stack need, pointer jumping for subtree
starts).  Run on the host lane by lane, it must produce the serial
compiler's bytecode instruction for instruction (the serial compiler is the
one tests/test_bytecode.py pins against the oracle): on GA populations, on
skeletons with OP_REF leaves and OP_STORE markers, on the deepest and widest
trees the format allows, and it must reject what the serial compiler
rejects."""
import numpy as np
import pytest

from symbolic_regression_gpu_amd.api import SrError, compile_expr, compile_skeleton
from tests.helpers import ALL_BINARY, ALL_UNARY, random_population

OP_REF, OP_STORE = -2, -3


def _arity(t):
    return 2 if 1 <= t <= 9 else (0 if t in (0, -1, -2) else 1)


def _sizes(tok):
    n = len(tok)
    size = [0] * n
    for i in range(n - 1, -1, -1):
        s, p = 1, i + 1
        for _ in range(_arity(tok[i])):
            s += size[p]
            p += size[p]
        size[i] = s
    return size


def _both(tok, val, nv):
    a = compile_skeleton(tok, val, nv, algorithm=0)
    b = compile_skeleton(tok, val, nv, algorithm=1)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2], (tok, a, b)
    return a


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_wave_compiler_equals_serial_on_populations(seed):
    pop = random_population(150, seed, nv=5, max_tokens=200, max_depth=12, unary=ALL_UNARY, binary=ALL_BINARY,
                            const_min=-3, const_max=3)
    for e in range(pop.E):
        tok, val = pop.expr(e)
        val = val.astype(np.float32)
        code, arg, need = _both(tok, val, 5)
        c2, a2, n2 = compile_expr(tok, val, nv=5)
        assert np.array_equal(code, c2) and np.array_equal(arg, a2) and need == n2


def _skeleton(tok, val, rng):
    """Replace some subtrees by OP_REF leaves and mark some operators with
    OP_STORE, as the planner's skeletons do."""
    tok, val = list(tok), list(val)
    out_t, out_v = [], []
    size = _sizes(tok)
    i = 0
    slot = 0
    while i < len(tok):
        if i > 0 and size[i] >= 2 and rng.random() < 0.15:
            out_t.append(OP_REF)
            out_v.append(float(rng.integers(0, 4000)))
            i += size[i]
            continue
        if _arity(tok[i]) > 0 and size[i] >= 2 and rng.random() < 0.2:
            out_t.append(OP_STORE)
            out_v.append(float(slot))
            slot += 1
        out_t.append(tok[i])
        out_v.append(val[i])
        i += 1
    if _arity(tok[0]) > 0 and rng.random() < 0.5:  # a launch-1 program starts with its column's OP_STORE
        out_t.insert(0, OP_STORE)
        out_v.insert(0, 123.0)
    return np.array(out_t, np.int32), np.array(out_v, np.float32)


@pytest.mark.parametrize("seed", [5, 6, 7])
def test_wave_compiler_equals_serial_on_skeletons(seed):
    rng = np.random.default_rng(seed)
    pop = random_population(120, seed, nv=3, max_tokens=250, max_depth=14, unary=ALL_UNARY, binary=ALL_BINARY)
    refs = 0
    for e in range(pop.E):
        tok, val = pop.expr(e)
        st, sv = _skeleton(tok, val, rng)
        code, arg, _ = _both(st, sv, 3)
        is_ref = ((code >> 8) & 7) == 3
        refs += int(np.sum(is_ref))
        assert np.sum(code == 31 | (2 << 13)) == np.sum(st == OP_STORE)  # STORE_CODE per OP_STORE
        # prefetch marks: 1..4 on the first four OP_REF operands in program order
        marks = (code >> 16) & 7
        assert list(marks[is_ref][:4]) == list(range(1, min(4, int(np.sum(is_ref))) + 1))
        assert not np.any(marks[~is_ref]) and not np.any(marks[is_ref][4:])
    assert refs > 0


def _chain_unary(n):
    return [10] * (n - 1) + [-1], [0.0] * (n - 1) + [0.0]


def _chain_binary(n_ops, left):
    # left-deep: op(op(op(x, x), x), x); right-deep: op(x, op(x, op(x, x)))
    if left:
        tok = [2] * n_ops + [-1] * (n_ops + 1)
    else:
        tok = []
        for _ in range(n_ops):
            tok += [4, -1]
        tok += [-1]
    return tok, [0.0] * len(tok)


def _balanced(depth):
    if depth == 0:
        return [-1], [1.0]
    t, v = _balanced(depth - 1)
    return [2] + t + t, [0.0] + v + v


@pytest.mark.parametrize("case", ["unary1023", "left511", "right511", "balanced9", "mixed"])
def test_wave_compiler_deep_and_wide_trees(case):
    if case == "unary1023":
        tok, val = _chain_unary(1024)
    elif case == "left511":
        tok, val = _chain_binary(511, True)
    elif case == "right511":
        tok, val = _chain_binary(511, False)
    elif case == "balanced9":
        tok, val = _balanced(9)  # 1023 tokens, stack need 8
    else:
        t1, v1 = _balanced(5)
        t2, v2 = _chain_unary(200)
        t3, v3 = _chain_binary(100, False)
        tok = [3, 1] + t1 + t2 + t3
        val = [0.0, 0.0] + v1 + v2 + v3
    code, arg, need = _both(np.array(tok, np.int32), np.array(val, np.float32), 2)
    if case == "balanced9":
        assert need == 8


@pytest.mark.parametrize("tok", [[1, 0], [0, 0], [10], [28, 0], [-5], [OP_STORE, -1], [3, OP_STORE, 0, -1],
                                 [1, -1, -1, -1]])
def test_wave_compiler_rejects_malformed(tok):
    val = [1.0] * len(tok)
    for alg in (0, 1):
        with pytest.raises(SrError):
            compile_skeleton(tok, val, 2, algorithm=alg)


@pytest.mark.parametrize("tok,val", [([-1], [1.0]), ([0], [2.5]), ([OP_REF], [17.0]), ([-1], [9.0])])
def test_wave_compiler_single_leaf(tok, val):
    code, arg, need = _both(np.array(tok, np.int32), np.array(val, np.float32), 3)
    assert len(code) == 1 and need == 0
