"""Adjoint (reverse-mode) dMSE/dconstants on the GPU against the oracle's FP64
forward-mode gradient with its error bound for an fp32 adjoint evaluation
(sro_mse_grad64_bound, pinned to torch FP64 autograd on the CPU), GPU central
differences against FP64 central differences, the torch.autograd.Function
shim, and an Adam fit.

Tolerances:
  * adjoint, per slot, on every expression the bound covers (no protected
    branch within rounding at any valid point): |g_gpu - g64| <= 2 bound;
  * GPU central differences (h = 1e-3) vs FP64 central differences at the same
    fp32-rounded points, per slot whose two perturbed points are stable: the
    two MSEs of a difference each carry at most bound * mse of error, plus the
    fp32 rounding of the stored MSEs, so
      |g_fd - g_fd64| <= 2 (bound + 2^-23) mse / (2h) + 1e-3 |g_fd64|.
FD and adjoint are not compared with each other: FD's O(h^2) truncation term
is unbounded for the curvature exp/div produce.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from symbolic_regression_gpu_amd import EVAL_BATCHED
from symbolic_regression_gpu_amd.autograd import ConstantFit
from tests.helpers import mse_error_bound, to_api, uniform_data
from tests.test_adjoint_cpu import BINARY_DOMAIN, UNARY_DOMAIN, c5_population, single_op
from tests import torch_ref

pytestmark = pytest.mark.gpu


def _adjoint_vs_bound(ctx, pop, slot, X, nv):
    """Runs the adjoint on the GPU; returns (covered, mse, grad, consts)."""
    N = X.shape[1]
    ctx.set_data(X)
    fit = ConstantFit(ctx, to_api(pop), slot)
    c0 = fit.initial()
    mse, g = fit.adjoint_grad(c0)
    torch.cuda.synchronize()
    mse, g, c0 = mse.cpu().numpy(), g.cpu().numpy(), c0.cpu().numpy()
    covered = 0
    for e in range(pop.E):
        tok, val = pop.expr(e)
        sl = slot[pop.off[e]:pop.off[e] + pop.len[e]]
        m64, g64, b64, cnt = O.mse_grad64_bound(tok, val, sl, X, nv, N, c0[e], max(N // 64, 1))
        if cnt == 0:
            assert np.isnan(mse[e])
            continue
        if not np.isfinite(b64).all():
            continue
        covered += 1
        assert (np.abs(g[e] - g64) <= 2 * b64).all(), (e, g[e], g64, b64)
    return covered, mse, g, c0, fit


@pytest.mark.parametrize("min_len,max_tokens", [(50, 200), (10, 60)])
def test_adjoint_matches_fp64_gradient_and_fd(ctx, min_len, max_tokens):
    nv, N = 4, 4096
    pop, slot = c5_population(48, 5, nv, min_len=min_len, max_tokens=max_tokens)
    X = uniform_data(nv, N, 3)
    X[nv] = np.sin(X[0]) * X[1] + 0.5 * X[2]
    covered, mse, g, c0, fit = _adjoint_vs_bound(ctx, pop, slot, X, nv)
    # forward MSE equals the plain evaluator's
    m_b, _, _ = ctx.eval_mse(to_api(pop), EVAL_BATCHED)
    fin = np.isfinite(m_b)
    assert np.array_equal(fin, np.isfinite(mse))
    assert np.allclose(mse[fin], m_b[fin], rtol=1e-6)
    g_fd = fit.fd_grad(torch.as_tensor(c0, device="cuda"), 1e-3)
    torch.cuda.synchronize()
    g_fd = g_fd.cpu().numpy()
    fd_checked = 0
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
        m64 = float(np.mean((v64[valid] - X[nv][valid]) ** 2))
        fd64 = torch_ref.fd_grad(tok, val, X[:nv], X[nv], c0[e], sl, valid, 1e-3)
        for k in np.unique(sl[sl >= 0]):
            # the bound must hold at both perturbed points, where a protected
            # branch may come within rounding even though c0 is clear of it
            rbk = 0.0
            for sgn in (1.0, -1.0):
                vk = val.copy()
                vk[sl == k] = np.float32(np.float32(val[sl == k][0]) + np.float32(sgn * 1e-3))
                pk = O.predict32(tok, vk, X, nv, N)
                vk64, bk64 = O.predict64_bound(tok, vk, X, nv, N)
                rbk = max(rbk, mse_error_bound(vk64, bk64, pk, X[nv]))
            if not np.isfinite(rbk):
                continue
            tol = 2 * (rbk + 2.0 ** -23) * m64 / 2e-3 + 1e-3 * abs(fd64[k])
            assert abs(g_fd[e][k] - fd64[k]) <= tol, (e, k, g_fd[e], fd64, tol)
            fd_checked += 1
    print(f"adjoint within the FP64 gradient bound on {covered} of {pop.E} expressions, "
          f"FD on {fd_checked} slots")
    assert covered >= 8 and fd_checked >= 10


@pytest.mark.parametrize("nv", [6, 8, 12, 20])
def test_adjoint_variable_classes(ctx, nv):
    """Each variable-register class of the constant-fitting kernels (<= 8:
    nine GPR-indexed registers, <= 16, <= 32): gradients within the FP64
    bound, forward MSEs (adjoint pass and forward-only pass) equal the
    batched evaluator's."""
    N = 4096
    pop, slot = c5_population(32, 40 + nv, nv, min_len=20, max_tokens=80)
    X = uniform_data(nv, N, 50 + nv)
    X[nv] = np.sin(X[0]) * X[nv - 1] + 0.5 * X[nv // 2]
    covered, mse, g, c0, fit = _adjoint_vs_bound(ctx, pop, slot, X, nv)
    assert covered >= pop.E // 4, covered
    m_b, _, _ = ctx.eval_mse(to_api(pop), EVAL_BATCHED)
    fin = np.isfinite(m_b)
    assert np.array_equal(fin, np.isfinite(mse))
    assert np.allclose(mse[fin], m_b[fin], rtol=1e-6)
    m_f = fit.mse(torch.as_tensor(c0, device="cuda")).detach().cpu().numpy()
    assert np.array_equal(np.isfinite(m_f), fin)
    assert np.allclose(m_f[fin], m_b[fin], rtol=1e-6)


def test_every_opcode_partials_on_gpu(ctx):
    from oracle.oracle import Population as OPop
    ops = sorted(UNARY_DOMAIN) + sorted(BINARY_DOMAIN)
    for op in ops:
        tok, val, sl, lo, hi = single_op(op)
        nv, N = 2, 1024
        X = uniform_data(nv, N, op, lo=lo, hi=hi)
        if op >= 10:
            X[0] = (X[0] - np.float32(0.05)) / np.float32(0.3)
        X[nv] = np.cos(X[0])
        pop = OPop([(tok, val)])
        covered, _, _, _, _ = _adjoint_vs_bound(ctx, pop, sl, X, nv)
        assert covered == 1, op


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


def test_param_autotune_keeps_gradients(ctx):
    nv, N = 4, 8192
    pop, slot = c5_population(32, 11, nv, min_len=20, max_tokens=80)
    X = uniform_data(nv, N, 12)
    X[nv] = np.sin(X[0]) + X[1]
    ctx.set_data(X)
    fit = ConstantFit(ctx, to_api(pop), slot)
    c0 = fit.initial()
    m1, g1 = fit.adjoint_grad(c0)
    f1 = fit.mse(c0.detach())
    fd, ad = ctx.param_autotune(c0.data_ptr())
    assert fd in (1, 2, 4) and ad in (1, 2)
    m2, g2 = fit.adjoint_grad(c0)
    f2 = fit.mse(c0.detach())
    torch.cuda.synchronize()
    # only the order of partial sums can move with the launch shape
    fin = torch.isfinite(m1)
    assert torch.equal(fin, torch.isfinite(m2))
    assert torch.allclose(m1[fin], m2[fin], rtol=1e-6) and torch.allclose(f1[fin], f2[fin], rtol=1e-6)
    a, b = torch.nan_to_num(g1[fin]).double(), torch.nan_to_num(g2[fin]).double()
    scale = a.abs().amax(dim=1, keepdim=True)
    assert ((a - b).abs() <= 1e-3 * a.abs() + 1e-5 * scale).all()
    print(f"param autotune: forward dpt {fd}, adjoint dpt {ad}")
