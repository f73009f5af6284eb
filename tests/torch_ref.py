"""FP64 torch evaluation of prefix expressions with autograd, used to check the
adjoint kernel's dMSE/dconstants.  Semantics follow the fp64 twin of the
reference evaluator (cpu_simple_single.cpp:6-79): MIN/MAX by <= / >=, DIV by 0
-> NaN, LOOSE_* guards with 1e-9, clamped exp (opcode 29) -> 1e9 above 10."""
from __future__ import annotations

import numpy as np
import torch

D, MAXV = 1e-9, 1e9


def _guard(x):
    return torch.where(x.abs() <= D, torch.where(x < 0, torch.full_like(x, -D), torch.full_like(x, D)), x)


def _unary(op, a):
    if op == 10: return torch.sin(a)
    if op == 11: return torch.cos(a)
    if op == 12: return torch.tan(a)
    if op == 13: return torch.sinh(a)
    if op == 14: return torch.cosh(a)
    if op == 15: return torch.tanh(a)
    if op == 16: return torch.exp(a)
    if op == 17: return torch.log(a)
    if op == 18: return 1.0 / a
    if op == 19: return torch.asin(a)
    if op == 20: return torch.acos(a)
    if op == 21: return torch.atan(a)
    if op == 22: return torch.where(a == 0, torch.full_like(a, -MAXV), torch.log(torch.where(a == 0, torch.ones_like(a), a.abs())))
    if op == 23: return 1.0 / _guard(a)
    if op == 24: return a.abs()
    if op == 25: return -a
    if op == 26: return torch.sqrt(a)
    if op == 27: return torch.sqrt(a.abs())
    if op == 29: return torch.where(a > 10, torch.full_like(a, MAXV), torch.exp(torch.clamp(a, max=10.0)))
    return torch.zeros_like(a)


def _binary(op, a, b):
    if op == 1: return a + b
    if op == 2: return a - b
    if op == 3: return a * b
    if op == 4: return torch.where(b == 0, torch.full_like(a, float("nan")), a / torch.where(b == 0, torch.ones_like(b), b))
    if op == 5: return torch.pow(a, b)
    if op == 6: return torch.where(a <= b, a, b)
    if op == 7: return torch.where(a >= b, a, b)
    if op == 8: return a / _guard(b)
    if op == 9:
        both0 = (a == 0) & (b == 0)
        return torch.where(both0, torch.zeros_like(a), torch.pow(torch.where(both0, torch.ones_like(a), a.abs()), b))
    return torch.zeros_like(a)


def eval_expr(tok, val, X64, consts, const_slot):
    """tok/val: one expression; X64: [nv, N] float64; consts: float64 tensor
    [8] (free constants, requires_grad); const_slot: per token slot or -1."""
    N = X64.shape[1]
    stack = []
    for i in range(len(tok) - 1, -1, -1):
        t = int(tok[i])
        if t == 0:
            s = int(const_slot[i])
            c = consts[s] if s >= 0 else torch.tensor(float(np.float32(val[i])), dtype=torch.float64)
            stack.append(c.expand(N))
        elif t == -1:
            k = int(val[i])
            stack.append(X64[k] if k < X64.shape[0] else torch.zeros(N, dtype=torch.float64))
        elif 1 <= t <= 9:
            a = stack.pop()
            b = stack.pop()
            stack.append(_binary(t, a, b))
        else:
            stack.append(_unary(t, stack.pop()))
    return stack[0]


def mse_and_grad(tok, val, X32, y32, cvals8, const_slot, valid_mask):
    """FP64 MSE over valid_mask points and its gradient wrt the 8 slots."""
    X64 = torch.tensor(X32, dtype=torch.float64)
    y64 = torch.tensor(y32, dtype=torch.float64)
    c = torch.tensor(np.asarray(cvals8, np.float64), requires_grad=True)
    pred = eval_expr(tok, val, X64, c, const_slot)
    m = torch.tensor(valid_mask)
    if m.sum() == 0:
        return float("nan"), np.zeros(8)
    loss = ((pred[m] - y64[m]) ** 2).mean()
    loss.backward()
    g = c.grad.numpy() if c.grad is not None else np.zeros(8)
    return float(loss.detach()), np.nan_to_num(g)


def fd_grad(tok, val, X32, y32, cvals8, const_slot, valid_mask, h):
    """FP64 central differences (m(c_hi) - m(c_lo)) / (c_hi - c_lo) over the
    same valid points, for the 8 slots, where c_hi/c_lo are c_k +- h rounded
    to fp32 (the constants the fp32 replicas hold)."""
    X64 = torch.tensor(X32, dtype=torch.float64)
    y64 = torch.tensor(y32, dtype=torch.float64)
    m = torch.tensor(valid_mask)
    base = np.asarray(cvals8, np.float64)
    g = np.zeros(8)
    used = set(int(s) for s in const_slot if s >= 0)
    with torch.no_grad():
        for k in sorted(used):
            out, at = [], []
            for sgn in (1.0, -1.0):
                c = base.copy()
                c[k] = np.float64(np.float32(np.float32(c[k]) + np.float32(sgn * h)))
                pred = eval_expr(tok, val, X64, torch.tensor(c), const_slot)
                out.append(float(((pred[m] - y64[m]) ** 2).mean()))
                at.append(c[k])
            g[k] = (out[0] - out[1]) / (at[0] - at[1])
    return g
