"""Shared test helpers: seeded populations from the oracle's generator, the
parity tolerance, and a bytecode emulator that checks the batched
interpreter's compiled programs against the oracle on the CPU."""
from __future__ import annotations

import numpy as np

from oracle import oracle as O

# The bench grammar: binary {+,-,*,protected /}, unary {sin, cos, exp(clamped),
# protected log, sqrt(abs)}.
GRAMMAR_UNARY = [10, 11, 29, 22, 27]
GRAMMAR_BINARY = [1, 2, 3, 8]
ALL_UNARY = list(range(10, 28)) + [29]
ALL_BINARY = list(range(1, 10))

EPS32 = float(np.finfo(np.float32).eps)


def random_population(n, seed, nv=3, max_tokens=60, max_depth=6, unary=None, binary=None,
                      const_min=-2.0, const_max=2.0, min_len=1):
    """n expressions from the oracle's Philox generator (utils.cpp grammar)."""
    p = O.ga_params(num_vars=nv, max_tokens=max_tokens, max_depth=max_depth,
                    const_min=const_min, const_max=const_max,
                    unary_ops=unary or ALL_UNARY, binary_ops=binary or ALL_BINARY)
    cap = max_tokens + max_depth + 2
    exprs = []
    s = seed
    while len(exprs) < n:
        k = max(2 * (n - len(exprs)), 16)
        rc, tok, val, ln = O.ga_init_population(p, k, cap, s)
        assert rc == 0
        for i in range(k):
            if ln[i] >= min_len and len(exprs) < n:
                exprs.append((tok[i * cap:i * cap + ln[i]].copy(),
                              val[i * cap:i * cap + ln[i]].astype(np.float64)))
        s += 1
    return O.Population(exprs)


def planted_population(n, seed, nv=3, pool_size=10, unary=None, binary=None):
    """n expressions sharing subtrees: about half join two members of a pool
    of random subtrees (2..15 tokens) under a binary operator, a third wrap a
    pool member and a fresh expression, the rest are fresh; every pool member
    recurs many times."""
    rng = np.random.default_rng(seed)
    ub, bb = unary or ALL_UNARY, binary or ALL_BINARY
    base = random_population(pool_size + n, seed, nv=nv, max_tokens=15, max_depth=4, unary=ub, binary=bb,
                             min_len=2)
    pool = [base.expr(i) for i in range(pool_size)]
    exprs = []
    for i in range(n):
        r = rng.random()
        a = pool[rng.integers(pool_size)]
        if r < 0.5:
            b = pool[rng.integers(pool_size)]
        elif r < 0.8:
            b = base.expr(pool_size + i)
        else:
            exprs.append(base.expr(pool_size + i))
            continue
        op = bb[rng.integers(len(bb))]
        if rng.random() < 0.5:
            a, b = b, a
        t = np.concatenate([[op], a[0], b[0]]).astype(np.int32)
        v = np.concatenate([[0.0], a[1], b[1]])
        if rng.random() < 0.3:
            t = np.concatenate([[ub[rng.integers(len(ub))]], t]).astype(np.int32)
            v = np.concatenate([[0.0], v])
        exprs.append((t, v))
    return O.Population(exprs)


def collision_population():
    """x0*x0 and x2*x2 (VAR values 0.0 and 2.0), six times each as
    sin(square) + c: the flat FNV hash of the two subtrees is equal, so a cache
    keyed by that hash alone would hand one the other's column."""
    sq0 = ([3, -1, -1], [0.0, 0.0, 0.0])
    sq2 = ([3, -1, -1], [0.0, 2.0, 2.0])
    exprs = []
    for k in range(6):
        for sq in (sq0, sq2):
            exprs.append((np.array([1, 10] + sq[0] + [0], np.int32),
                          np.array([0.0, 0.0] + sq[1] + [0.25 * k + 0.5], np.float64)))
    return O.Population(exprs)


def to_api(pop):
    from symbolic_regression_gpu_amd import Population
    return Population(pop.tok, pop.val, pop.off, pop.len)


def uniform_data(nv, N, seed, lo=1.0, hi=5.0):
    rng = np.random.default_rng(seed)
    return rng.uniform(lo, hi, size=(nv + 1, N)).astype(np.float32)


def ulp(x):
    x = np.abs(np.asarray(x, np.float32))
    return np.spacing(x).astype(np.float64)


def point_parity(gpu, cpu32, v64, b64, k=2.0):
    """Per-point fp32 parity against the FP64 running-error bound
    (oracle sro_predict64_bound): a point whose bound is infinite (a protected
    branch or the fp32 range can flip within rounding) is accepted; otherwise
    the finite / NaN / Inf class must equal the CPU oracle's and
        |gpu - v64| <= k * max(bound, |cpu32 - v64|) + 2 ulp(v64).
    Returns a bool mask."""
    g = np.asarray(gpu, np.float64)
    c = np.asarray(cpu32, np.float64)
    v = np.asarray(v64, np.float64)
    b = np.asarray(b64, np.float64)
    stable = np.isfinite(b)
    same_nan = np.isnan(g) & np.isnan(c)
    same_inf = np.isinf(g) & np.isinf(c) & (np.sign(g) == np.sign(c))
    with np.errstate(invalid="ignore", over="ignore"):
        err = np.where(np.isfinite(c) & np.isfinite(v), np.abs(c - v), 0.0)
        tol = k * np.maximum(b, err) + 2.0 * ulp(np.where(np.isfinite(v), v, 0.0))
        close = np.isfinite(g) & np.isfinite(c) & (np.abs(g - v) <= tol)
    return ~stable | same_nan | same_inf | close


def mse_error_bound(v64, b64, c32, y):
    """Bound on |MSE_fp32 - MSE_exact| over the finite points, or inf when a
    point is unstable.  Relative to the FP64 MSE."""
    v = np.asarray(v64, np.float64)
    b = np.asarray(b64, np.float64)
    fin = np.isfinite(np.asarray(c32, np.float64))
    if not np.isfinite(b[fin]).all() or fin.sum() == 0:
        return np.inf
    e = np.maximum(b, np.abs(np.asarray(c32, np.float64) - v))[fin]
    d = np.abs(v[fin] - np.asarray(y, np.float64)[fin])
    mse = np.mean(d * d)
    bound = np.mean(2 * d * e + e * e) + 2 * EPS32 * mse
    return bound / max(mse, 1e-300)


def check_population(pop, X, nv, N, m_gpu, v_gpu, pred_gpu=None, rtol=1e-5, floor_mse=0.8, floor_pts=0.9,
                     name=""):
    """Conditioning-aware parity of a GPU evaluation (MSE, valid counts and
    optionally predictions [E, N]) with the oracle.

    Every expression: the valid count differs from the oracle's by at most
    the number of points whose FP64 error bound is infinite (points that can
    change class within rounding), so it is equal where there are none; the
    MSE is NaN exactly where the oracle's is unless no valid point is pinned
    that way (an MSE is NaN only when its valid count is 0); the MSE lies
    within twice its bound of the FP64 MSE.  Expressions whose relative MSE error bound is
    <= 1e-6 must match the oracle's MSE within rtol.  Predictions must pass
    point_parity.

    The bars as stated for the evaluator -- MSE within 1e-5 relative of the
    oracle, per-point outputs within 4 ulp (of the oracle's fp32 value) -- are
    measured over the whole population: the shares that meet them must reach
    floor_mse / floor_pts, and the worst and median relative MSE error against
    the FP64 evaluator are printed.  Returns those statistics (dict)."""
    m_gpu = np.asarray(m_gpu, np.float32)
    m_ref, v_ref = O.mse32(pop, X, nv, N, threads=8)
    y = X[nv]
    strict = 0
    pts_ok = pts_n = 0
    e64 = []
    for e in range(pop.E):
        tok, val = pop.expr(e)
        c32 = O.predict32(tok, val, X, nv, N)
        v64, b64 = O.predict64_bound(tok, val, X, nv, N)
        rb = mse_error_bound(v64, b64, c32, y)
        if pred_gpu is not None:
            g = np.asarray(pred_gpu[e], np.float32)
            ok = point_parity(g, c32, v64, b64)
            assert ok.all(), (e, np.flatnonzero(~ok)[:5])
            same_class = (np.isnan(g) & np.isnan(c32)) | (np.isinf(g) & (g == c32))
            fin = np.isfinite(g) & np.isfinite(c32)
            with np.errstate(invalid="ignore", over="ignore"):
                near = fin & (np.abs(g.astype(np.float64) - c32) <= 4 * ulp(c32))
            pts_ok += int((same_class | near).sum())
            pts_n += N
        unstable = int((~np.isfinite(b64)).sum())
        assert abs(int(v_gpu[e]) - int(v_ref[e])) <= unstable, (e, v_gpu[e], v_ref[e], unstable)
        assert np.isnan(m_gpu[e]) == np.isnan(m_ref[e]) or unstable > 0, (e, m_gpu[e], m_ref[e])
        if np.isfinite(rb):
            assert v_gpu[e] == v_ref[e], (e, v_gpu[e], v_ref[e])
        if rb <= 1e-6:
            strict += 1
            assert mse_close([m_gpu[e]], [m_ref[e]], rtol).all(), (e, m_gpu[e], m_ref[e])
        fin = np.isfinite(c32)
        if fin.any() and np.isfinite(m_gpu[e]):
            m64 = np.mean((v64[fin] - y[fin]) ** 2)
            if np.isfinite(m64) and m64 > 0:
                e64.append(abs(float(m_gpu[e]) - m64) / m64)
            if np.isfinite(rb):
                assert abs(m_gpu[e] - m64) <= 2 * rb * m64 + 4 * EPS32 * m64, (e, m_gpu[e], m64, rb)
    both = np.isfinite(m_gpu) & np.isfinite(m_ref)
    within = mse_close(m_gpu[both], m_ref[both], 1e-5)
    stats = {
        "exprs": pop.E,
        "strict": strict,
        "mse_share_1e-5": float(within.mean()) if both.any() else 1.0,
        "points_share_4ulp": pts_ok / pts_n if pts_n else None,
        "max_rel_err_fp64": float(np.max(e64)) if e64 else 0.0,
        "median_rel_err_fp64": float(np.median(e64)) if e64 else 0.0,
    }
    print(f"[parity {name}] {stats}")
    assert stats["mse_share_1e-5"] >= floor_mse, stats
    if pts_n:
        assert stats["points_share_4ulp"] >= floor_pts, stats
    return stats


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.abs(a - b) / np.maximum(np.abs(b), 1e-30)


def mse_close(gpu, cpu, rtol=1e-5):
    gpu = np.asarray(gpu, np.float64)
    cpu = np.asarray(cpu, np.float64)
    nan_ok = np.isnan(gpu) == np.isnan(cpu)
    inf_ok = np.isinf(gpu) == np.isinf(cpu)
    fin = np.isfinite(gpu) & np.isfinite(cpu)
    close = np.ones_like(fin)
    close[fin] = rel_err(gpu[fin], cpu[fin]) <= rtol
    return nan_ok & inf_ok & close


# ---------------------------------------------------------------------------
# Bytecode emulator (CPU): runs a program from sr_compile_expr with the
# oracle's fp32 op table, one datapoint at a time.
# ---------------------------------------------------------------------------
SRC_NONE, SRC_CONST, SRC_VAR, SRC_REF, SRC_POP = 0, 1, 2, 3, 4
BC_PUSH, BC_SWAP = 1 << 11, 1 << 12


def run_bytecode(code, arg, x, cache_row=None):
    stk = []
    top = np.float32(0)
    o = np.float32(0)
    for c, a in zip(code.tolist(), arg.tolist()):
        if c & BC_PUSH:
            stk.append(top)
        src = (c >> 8) & 7
        op = c & 0xFF
        if src == SRC_CONST:
            o = np.array([a], np.int32).view(np.float32)[0]
        elif src == SRC_VAR:
            o = np.float32(x[a])
        elif src == SRC_REF:
            o = np.float32(cache_row[a])
        elif src == SRC_POP:
            o = stk.pop()
        if op == 0:
            top = o
        elif op < 10:
            if c & BC_SWAP:
                top, o = o, top
            top = np.float32(O.op32(op, float(top), float(o)))
        else:
            if src != SRC_NONE:
                top = o
            top = np.float32(O.op32(op, float(top), 0.0))
    return top, len(stk)
