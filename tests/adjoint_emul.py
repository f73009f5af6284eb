"""CPU emulation (numpy fp32, vectorised over datapoints) of the adjoint pass
the GPU runs on a parametric program from sr_compile_param_expr: a forward
sweep that records the accumulator after every instruction, then a reverse
sweep with an adjoint stack mirroring the value stack.  It checks the program
format (free-constant slot bits, POP producer indices, swap flags) and the
reverse-sweep bookkeeping on the CPU against FP64 autograd; the GPU kernel is
checked against the same references on the GPU."""
from __future__ import annotations

import numpy as np

from symbolic_regression_gpu_amd.api import compile_expr

f32 = np.float32
DELTA = f32(1e-9)
SRC_NONE, SRC_CONST, SRC_VAR, SRC_POP = 0, 1, 2, 4
BC_PUSH, BC_SWAP = 1 << 11, 1 << 12


def _guard(b):
    return np.where(np.abs(b) <= DELTA, np.where(b < 0, -DELTA, DELTA), b).astype(f32)


def unary(op, u):
    with np.errstate(all="ignore"):
        r = {
            10: lambda: np.sin(u), 11: lambda: np.cos(u), 12: lambda: np.tan(u),
            13: lambda: np.sinh(u), 14: lambda: np.cosh(u), 15: lambda: np.tanh(u),
            16: lambda: np.exp(u), 17: lambda: np.log(u), 18: lambda: f32(1) / u,
            19: lambda: np.arcsin(u), 20: lambda: np.arccos(u), 21: lambda: np.arctan(u),
            22: lambda: np.where(u == 0, f32(-1e9), np.log(np.abs(u))),
            23: lambda: f32(1) / _guard(u), 24: lambda: np.abs(u), 25: lambda: -u,
            26: lambda: np.sqrt(u), 27: lambda: np.sqrt(np.abs(u)),
            29: lambda: np.where(u > 10, f32(1e9), np.exp(np.minimum(u, f32(10)))),
        }[op]()
    return np.asarray(r, f32)


def binary(op, a, b):
    with np.errstate(all="ignore"):
        r = {
            1: lambda: a + b, 2: lambda: a - b, 3: lambda: a * b,
            4: lambda: np.where(b == 0, f32(np.nan), a / np.where(b == 0, f32(1), b)),
            5: lambda: np.power(a, b), 6: lambda: np.fmin(a, b), 7: lambda: np.fmax(a, b),
            8: lambda: a / _guard(b),
            9: lambda: np.where((a == 0) & (b == 0), f32(0), np.power(np.abs(a), b)),
        }[op]()
    return np.asarray(r, f32)


def d_unary(op, u, y):
    one = f32(1)
    with np.errstate(all="ignore"):
        r = {
            10: lambda: np.cos(u), 11: lambda: -np.sin(u), 12: lambda: one + y * y,
            13: lambda: np.cosh(u), 14: lambda: np.sinh(u), 15: lambda: one - y * y,
            16: lambda: y, 17: lambda: one / u, 18: lambda: -(y * y),
            19: lambda: one / np.sqrt(one - u * u), 20: lambda: -one / np.sqrt(one - u * u),
            21: lambda: one / (one + u * u),
            22: lambda: np.where(u == 0, f32(0), one / u),
            23: lambda: np.where(np.abs(u) <= DELTA, f32(0), -(y * y)),
            24: lambda: np.sign(u), 25: lambda: np.full_like(u, -1),
            26: lambda: f32(0.5) / y,
            27: lambda: np.where(u == 0, f32(0), np.where(u > 0, f32(0.5), f32(-0.5)) / y),
            29: lambda: np.where(u > 10, f32(0), y),
        }[op]()
    return np.asarray(r, f32)


def d_binary(op, a, b, y):
    one = f32(1)
    with np.errstate(all="ignore"):
        if op == 1:
            return np.ones_like(a), np.ones_like(a)
        if op == 2:
            return np.ones_like(a), -np.ones_like(a)
        if op == 3:
            return b, a
        if op == 4:
            return (one / b).astype(f32), (-y / b).astype(f32)
        if op == 5:
            da = np.where(a == 0, np.where(b == 1, one, f32(0)), b * y / a)
            db = np.where(a > 0, y * np.log(np.where(a > 0, a, one)), f32(0))
            return da.astype(f32), db.astype(f32)
        if op in (6, 7):
            cmp = (a <= b) if op == 6 else (a >= b)
            sel = np.isnan(b) | (~np.isnan(a) & cmp)
            return sel.astype(f32), (~sel).astype(f32)
        if op == 8:
            small = np.abs(b) <= DELTA
            d = _guard(b)
            return (one / d).astype(f32), np.where(small, f32(0), -y / d).astype(f32)
        if op == 9:
            z = a == 0
            da = np.where(z, f32(0), b * y / np.where(z, one, a))
            db = np.where(z, f32(0), y * np.log(np.abs(np.where(z, one, a))))
            return da.astype(f32), db.astype(f32)
    raise ValueError(op)


BC_ACTIVE = 1 << 15
BC_TAPE_SHIFT = 20
CL_LOAD, CL_BIN, CL_UN = 0, 1, 2


def mse_and_grad(tok, val, slot, X, nv, cvals):
    """(mse, grad[8]) of one expression over X [nv+1, N] (y last) the way the
    adjoint kernel computes them: forward pass taping the rows the program
    asks for, reverse pass over the active instructions only (jumping over
    subtrees without a free constant), fp32 values and adjoints, per-point
    gradients summed in fp64 here (the kernel sums in fp32 per lane)."""
    code, arg, _, aux, rows = compile_expr(tok, val, nv, slot)
    N = X.shape[1]
    xs = [X[k].astype(f32) for k in range(nv)]

    def cval(c, a):
        k = ((c >> 16) & 15) - 1
        return cvals[k] if k >= 0 else np.array([a], np.int32).view(f32)[0]

    n = len(code)
    tape = np.zeros((max(rows, 1), N), f32)
    stk = []
    top = np.zeros(N, f32)
    for i in range(n):
        c, a = int(code[i]), int(arg[i])
        s, op, cls = (c >> 8) & 7, c & 0xFF, (c >> 13) & 3
        if c & BC_PUSH:
            stk.append(top)
        if s == SRC_POP:
            o = stk.pop()
        elif s == SRC_VAR:
            o = xs[a] if a < nv else np.zeros(N, f32)
        else:
            o = np.full(N, cval(c, a), f32)
        if cls == CL_LOAD:
            top = o
        elif cls == CL_BIN:
            A, B = (o, top) if c & BC_SWAP else (top, o)
            top = binary(op, A, B)
        else:
            top = unary(op, o if s != SRC_NONE else top)
        r = ((c >> BC_TAPE_SHIFT) & 0x7FF) - 1
        if r >= 0:
            tape[r] = top
    y = X[nv].astype(f32)
    ok = np.isfinite(top)
    diff = (top - y).astype(f32)
    at = np.where(ok, diff, f32(0)).astype(f32)
    astk = []
    tg = np.zeros((8, N), f32)
    i = n - 1
    while i >= 0:
        c, a, x = int(code[i]), int(arg[i]), int(aux[i])
        if not c & BC_ACTIVE:
            at = astk.pop() if code[x] & BC_PUSH else np.zeros(N, f32)
            i = x - 1
            continue
        s, op, cls = (c >> 8) & 7, c & 0xFF, (c >> 13) & 3
        yo = tape[((c >> BC_TAPE_SHIFT) & 0x7FF) - 1]
        x_in = tape[x] if x >= 0 else np.zeros(N, f32)
        if s == SRC_POP:
            ov = tape[a]
        elif s == SRC_VAR:
            ov = xs[a] if a < nv else np.zeros(N, f32)
        else:
            ov = np.full(N, cval(c, a), f32)
        a_o = np.zeros(N, f32)
        a_in = np.zeros(N, f32)
        if cls == CL_LOAD:
            a_o = at
        elif cls == CL_BIN:
            sw = bool(c & BC_SWAP)
            A, B = (ov, x_in) if sw else (x_in, ov)
            dA, dB = d_binary(op, A, B, yo)
            gA, gB = (at * dA).astype(f32), (at * dB).astype(f32)
            a_o, a_in = (gA, gB) if sw else (gB, gA)
        else:
            g = (at * d_unary(op, ov if s != SRC_NONE else x_in, yo)).astype(f32)
            if s != SRC_NONE:
                a_o = g
            else:
                a_in = g
        k = ((c >> 16) & 15) - 1
        if s == SRC_CONST and k >= 0:
            tg[k] += a_o
        elif s == SRC_POP:
            astk.append(a_o)
        if c & BC_PUSH:
            a_in = astk.pop()
        at = a_in
        i -= 1
    cnt = int(ok.sum())
    if cnt == 0:
        return float("nan"), np.full(8, np.nan)
    mse = float(np.sum(diff[ok].astype(np.float64) ** 2) / cnt)
    return mse, 2.0 * np.where(ok, tg, 0).astype(np.float64).sum(axis=1) / cnt
