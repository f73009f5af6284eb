"""Adjoint (reverse-mode) dMSE/dconstants against FP64 torch autograd and
central finite differences (h = 1e-3), plus the torch.autograd.Function shim
and an Adam fit.

Tolerances: adjoint vs FP64 autograd, per expression,
  ||g_adj - g_64|| <= 2e-3 * ||g_64|| + 1e-6 * (1 + mse)
on expressions whose fp32 predictions are faithful everywhere (see
helpers.mse_error_bound <= 1e-4); vs GPU central differences (which carry O(h^2)
truncation and fp32 rounding of the MSE difference),
  ||g_adj - g_fd|| <= 5e-2 * ||g_adj|| + 2e-3 * (1 + mse).
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from symbolic_regression_gpu_amd import EVAL_BATCHED
from symbolic_regression_gpu_amd.autograd import ConstantFit
from tests.helpers import (GRAMMAR_BINARY, GRAMMAR_UNARY, mse_error_bound, random_population,
                           to_api, uniform_data)
from tests import torch_ref

pytestmark = pytest.mark.gpu


def c5_population(E, seed, nv, min_len=50, max_tokens=200):
    """C5 shape: E expressions of min_len..max_tokens tokens with 2..8 free
    constants (the first constants in token order; their values ~ N(0,1))."""
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


def test_adjoint_matches_fp64_autograd_and_fd(ctx):
    nv, N = 4, 4096
    pop, slot = c5_population(48, 5, nv)
    X = uniform_data(nv, N, 3)
    X[nv] = np.sin(X[0]) * X[1] + 0.5 * X[2]
    ctx.set_data(X)
    fit = ConstantFit(ctx, to_api(pop), slot)
    c0 = fit.initial()
    mse, g = fit.adjoint_grad(c0)
    g_fd = fit.fd_grad(c0, 1e-3)
    torch.cuda.synchronize()
    mse, g, g_fd = mse.cpu().numpy(), g.cpu().numpy(), g_fd.cpu().numpy()
    # forward MSE equals the plain evaluator's
    m_b, _, _ = ctx.eval_mse(to_api(pop), EVAL_BATCHED)
    fin = np.isfinite(m_b)
    assert np.array_equal(fin, np.isfinite(mse))
    assert np.allclose(mse[fin], m_b[fin], rtol=1e-6)
    checked = 0
    for e in range(pop.E):
        if not np.isfinite(mse[e]):
            continue
        tok, val = pop.expr(e)
        sl = slot[pop.off[e]:pop.off[e] + pop.len[e]]
        c32 = O.predict32(tok, val, X, nv, N)
        v64, b64 = O.predict64_bound(tok, val, X, nv, N)
        if not mse_error_bound(v64, b64, c32, X[nv]) <= 1e-4:
            continue
        valid = np.isfinite(c32)
        m64, g64 = torch_ref.mse_and_grad(tok, val, X[:nv], X[nv], c0[e].cpu().numpy(), sl, valid)
        err = np.linalg.norm(g[e] - g64)
        assert err <= 2e-3 * np.linalg.norm(g64) + 1e-6 * (1 + m64), (e, g[e], g64)
        err_fd = np.linalg.norm(g[e] - g_fd[e])
        assert err_fd <= 5e-2 * np.linalg.norm(g[e]) + 2e-3 * (1 + mse[e]), (e, g[e], g_fd[e])
        checked += 1
    print(f"adjoint checked against FP64 autograd on {checked} of {pop.E} expressions")
    assert checked >= pop.E // 3


def test_autograd_shim_and_adam(ctx):
    nv, N = 3, 8192
    pop, slot = c5_population(32, 9, nv, min_len=20, max_tokens=60)
    X = uniform_data(nv, N, 4)
    X[nv] = 1.5 * X[0] - np.cos(X[1]) + np.float32(0.01) * np.random.default_rng(0).normal(size=N).astype(np.float32)
    ctx.set_data(X)
    fit = ConstantFit(ctx, to_api(pop), slot)
    consts = fit.initial().requires_grad_()
    mse0 = fit.mse(consts)
    mse0.nansum().backward()
    mse1, g1 = fit.adjoint_grad(consts)
    assert torch.allclose(consts.grad, torch.nan_to_num(g1), rtol=0, atol=0)
    opt = torch.optim.Adam([consts], lr=2e-2)
    first = None
    for step in range(20):
        opt.zero_grad()
        m = fit.mse(consts)
        if first is None:
            first = m.detach().clone()
        m.nansum().backward()
        opt.step()
    last = fit.mse(consts.detach()).detach()
    ok = torch.isfinite(first) & torch.isfinite(last)
    improved = (last[ok] <= first[ok] * (1 + 1e-6)).float().mean().item()
    print(f"Adam: {improved:.2f} of finite expressions improved; "
          f"median MSE {first[ok].median().item():.4g} -> {last[ok].median().item():.4g}")
    assert improved >= 0.8
    assert torch.isfinite(consts).all()
