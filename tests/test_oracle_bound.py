"""The oracle's FP64 running error bound (sro_predict64_bound), which the GPU
parity tests use to tell rounding from mistakes, checked on the CPU:
  * the fp32 oracle itself always lies within the bound (all 28 opcodes, two
    data ranges), so a GPU result outside it is not explained by rounding;
  * an error of a few ulp planted in the fp32 values is caught;
  * points where rounding can flip a protected branch are marked unstable,
    and exact arithmetic gets a zero-width bound."""
import numpy as np

from oracle import oracle as O
from tests.helpers import ALL_BINARY, ALL_UNARY, point_parity, random_population, uniform_data


def _population_points(seed, nv, lo):
    pop = random_population(150, seed, nv=nv, max_tokens=70, max_depth=8, unary=ALL_UNARY,
                            binary=ALL_BINARY, const_min=-3, const_max=3)
    X = uniform_data(nv, 1500, seed, lo=lo, hi=5.0)
    for e in range(pop.E):
        tok, val = pop.expr(e)
        yield O.predict32(tok, val, X, nv, X.shape[1]), O.predict64_bound(tok, val, X, nv, X.shape[1])


def test_fp32_oracle_within_bound():
    stable = total = 0
    for seed, nv, lo in [(1000, 1, 1.0), (1003, 10, -3.0), (1017, 3, -3.0)]:
        for c32, (v64, b64) in _population_points(seed, nv, lo):
            assert point_parity(c32, c32, v64, b64).all()
            stable += int(np.isfinite(b64).sum())
            total += b64.size
    assert stable > 0.3 * total         # all opcodes on signed data: many protected branches


def test_planted_ulp_error_is_caught():
    caught = 0
    for c32, (v64, b64) in _population_points(1003, 10, -3.0):
        ok = np.isfinite(b64) & np.isfinite(c32) & (np.abs(v64) > 1e-3) & (b64 < 1e-6 * np.abs(v64))
        if not ok.any():
            continue
        bad = c32.copy()
        bad[ok] = bad[ok] * np.float32(1 + 2e-5)          # ~170 ulp
        caught += int((~point_parity(bad, c32, v64, b64)[ok]).all())
    assert caught > 10


def test_exact_and_unstable_cases():
    X = np.array([[2.0, 3.0], [0.0, 0.0]], np.float32)           # one variable, y unused
    # x + 1: exact in fp32 -> bound is the half-ulp of one rounding only
    v, b = O.predict64_bound(np.array([1, -1, 0], np.int32), np.array([0, 0, 1.0]), X, 1, 2)
    assert np.array_equal(v, [3.0, 4.0]) and (b <= 0.5 * 2 ** -23 * 4 + 1e-44).all()
    # log(sin(x) - sin(x)): the FP64 difference is exactly 0 but carries rounding
    # of sin -> sign unknown -> unstable
    tok = np.array([17, 2, 10, -1, 10, -1], np.int32)
    v, b = O.predict64_bound(tok, np.zeros(6), X, 1, 2)
    assert np.isinf(b).all()
    # protected log of an exact zero is exact (-1e9)
    tok = np.array([22, 2, -1, -1], np.int32)
    v, b = O.predict64_bound(tok, np.zeros(4), X, 1, 2)
    assert np.array_equal(v, [-1e9, -1e9]) and np.isfinite(b).all()
    # (-x) ^ (sin(x)): negative base with an inexact exponent -> NaN may flip
    tok = np.array([5, 25, -1, 10, -1], np.int32)
    v, b = O.predict64_bound(tok, np.zeros(5), X, 1, 2)
    assert np.isinf(b).all()
    # clamped exp right at its threshold after rounding: x * 5 with x = 2 -> 10
    tok = np.array([29, 3, -1, 0], np.int32)
    v, b = O.predict64_bound(tok, np.array([0, 0, 0, 5.0]), X[:, :1], 1, 1)
    assert v[0] == np.exp(10.0) and np.isfinite(b[0])
    tok = np.array([29, 3, 10, -1, 0], np.int32)
    Xs = np.array([[np.arcsin(10 / 10.5)], [0.0]], np.float32)
    v, b = O.predict64_bound(tok, np.array([0, 0, 0, 0, 10.5]), Xs, 1, 1)
    assert abs(v[0] - np.exp(10.0)) / np.exp(10.0) < 1e-5 and np.isinf(b[0])
