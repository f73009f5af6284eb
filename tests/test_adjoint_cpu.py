"""Parametric programs and the reverse sweep, on the CPU.

References: the oracle's forward-mode FP64 gradient of the MSE with an error
bound for an fp32 adjoint evaluation (sro_mse_grad64_bound), itself checked
against torch FP64 autograd; and the fp32 emulation of the GPU's adjoint pass
over programs from sr_compile_param_expr (tests/adjoint_emul.py).

  * FP64 oracle gradient == torch autograd (to 1e-3 of the fp32 bound);
  * emulated fp32 adjoint within 2x the bound on every expression the bound
    covers (no protected branch within rounding at any valid point);
  * every opcode's partials, alone, over a clean domain;
  * a wrong partial planted in the emulation is caught.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import adjoint_emul as AE
from tests import torch_ref
from tests.helpers import GRAMMAR_BINARY, GRAMMAR_UNARY, random_population, uniform_data


def c5_population(E, seed, nv, min_len=50, max_tokens=200):
    """E expressions with 2..8 free constants (the first constants in token
    order, values ~ N(0, 1))."""
    pop = random_population(E, seed, nv=nv, max_tokens=max_tokens, max_depth=10,
                            unary=GRAMMAR_UNARY, binary=GRAMMAR_BINARY, min_len=min_len)
    rng = np.random.default_rng(seed)
    slot = np.full(len(pop.tok), -1, np.int32)
    val = pop.val.copy()
    for e in range(pop.E):
        a, n = int(pop.off[e]), int(pop.len[e])
        consts = [i for i in range(a, a + n) if pop.tok[i] == 0]
        k = min(len(consts), int(rng.integers(2, 9)))
        for s, i in enumerate(consts[:k]):
            slot[i] = s
            val[i] = np.float32(rng.normal())
    pop.val = val
    return pop, slot


def slot_values(val, sl):
    c = np.zeros(8, np.float32)
    for v, k in zip(val, sl):
        if k >= 0:
            c[k] = np.float32(v)
    return c


def check_expr(tok, val, sl, X, nv, emul=AE.mse_and_grad):
    """Returns None when the bound does not cover the expression, else the
    largest |g_emul - g64| / bound over the slots (asserts the references)."""
    N = X.shape[1]
    c = slot_values(val, sl)
    m64, g64, b64, cnt = O.mse_grad64_bound(tok, val, sl, X, nv, N, c, max(N // 64, 1))
    if cnt == 0 or not np.isfinite(b64).all():
        return None
    valid = np.isfinite(O.predict32(tok, val, X, nv, N))
    mt, gt = torch_ref.mse_and_grad(tok, val, X[:nv], X[nv], c, sl, valid)
    assert abs(m64 - mt) <= 1e-9 * abs(mt)
    assert (np.abs(g64 - gt) <= 1e-3 * b64 + 1e-9 * np.abs(gt)).all(), (g64, gt, b64)
    _, g = emul(tok, val, sl, X, nv, c)
    return float(np.max(np.abs(g - g64) / np.maximum(b64, 1e-300)))


@pytest.mark.parametrize("min_len,max_tokens", [(50, 200), (10, 60)])
def test_emulated_adjoint_within_bound(min_len, max_tokens):
    nv, N = 4, 2048
    pop, slot = c5_population(40, 5, nv, min_len=min_len, max_tokens=max_tokens)
    X = uniform_data(nv, N, 3)
    X[nv] = np.sin(X[0]) * X[1] + 0.5 * X[2]
    covered = 0
    for e in range(pop.E):
        tok, val = pop.expr(e)
        sl = slot[pop.off[e]:pop.off[e] + pop.len[e]]
        r = check_expr(tok, val, sl, X, nv)
        if r is None:
            continue
        covered += 1
        assert r <= 2.0, (e, r)
    assert covered >= 8


# (expression template, data range) per opcode: op applied to c0 * x + c1
# (unary) or to (c0 * x, c1 + x) (binary), on a range clear of its branches.
UNARY_DOMAIN = {10: (1, 3), 11: (1, 3), 12: (0.1, 0.5), 13: (0.5, 2), 14: (0.5, 2), 15: (0.1, 1),
                16: (0.5, 2), 17: (1, 3), 18: (1, 3), 19: (0.1, 0.3), 20: (0.1, 0.3),
                21: (0.5, 2), 22: (1, 3), 23: (1, 3), 24: (1, 3), 25: (1, 3), 26: (1, 3),
                27: (1, 3), 29: (0.5, 2)}
BINARY_DOMAIN = {1: (1, 3), 2: (1, 3), 3: (1, 3), 4: (1, 3), 5: (1, 2), 6: (1, 3), 7: (1, 3),
                 8: (1, 3), 9: (1, 2)}


def single_op(op):
    # unary: op(c0 * x0 + c1) -> tokens [op, +, *, c0, x0, c1]
    if op >= 10:
        tok = np.array([op, 1, 3, 0, -1, 0], np.int32)
        val = np.array([0, 0, 0, 0.3, 0, 0.05])
        sl = np.array([-1, -1, -1, 0, -1, 1], np.int32)
        lo, hi = UNARY_DOMAIN[op]
    else:
        # binary: op(c0 * x0, c1 + x1) -> [op, *, c0, x0, +, c1, x1]
        tok = np.array([op, 3, 0, -1, 1, 0, -1], np.int32)
        val = np.array([0, 0, 1.1, 0, 0, 0.2, 1])
        sl = np.array([-1, -1, 0, -1, -1, 1, -1], np.int32)
        lo, hi = BINARY_DOMAIN[op]
    return tok, val, sl, lo, hi


@pytest.mark.parametrize("op", sorted(UNARY_DOMAIN) + sorted(BINARY_DOMAIN))
def test_every_opcode_partials(op):
    tok, val, sl, lo, hi = single_op(op)
    nv, N = 2, 1024
    X = uniform_data(nv, N, op, lo=lo, hi=hi)
    if op >= 10:
        X[0] = (X[0] - np.float32(0.05)) / np.float32(0.3)   # c0 * x + c1 spans [lo, hi]
    X[nv] = np.cos(X[0])
    r = check_expr(tok, val, sl, X, nv)
    assert r is not None and r <= 2.0, r


def test_planted_wrong_partial_is_caught():
    nv, N = 4, 2048
    pop, slot = c5_population(40, 5, nv, min_len=10, max_tokens=60)
    X = uniform_data(nv, N, 3)
    X[nv] = np.sin(X[0]) * X[1] + 0.5 * X[2]
    good = AE.d_unary

    def wrong(op, u, y):
        d = good(op, u, y)
        return -d if op == 10 else d                       # d sin(u) = -cos(u)

    caught = 0
    try:
        AE.d_unary = wrong
        for e in range(pop.E):
            tok, val = pop.expr(e)
            if 10 not in tok:
                continue
            sl = slot[pop.off[e]:pop.off[e] + pop.len[e]]
            r = check_expr(tok, val, sl, X, nv)
            caught += int(r is not None and r > 2.0)
    finally:
        AE.d_unary = good
    assert caught >= 3
