"""The two host-array entry points of the drop-in boundary that wrap the
device paths: sr_evolve_step (one select + vary step, cl_sr_system.cpp:285-330)
against the oracle's next generation bit for bit, and sr_grad_consts
(dMSE/dconstant per masked constant) against the oracle's FP64 gradient with
its fp32-adjoint error bound (|g - g64| <= 2 bound where the bound is
finite) and bit for bit against the constant-fitting path it wraps."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from symbolic_regression_gpu_amd.api import GaParams, Population, SelectParams, SrError
from symbolic_regression_gpu_amd.autograd import ConstantFit
from tests.helpers import GRAMMAR_BINARY, GRAMMAR_UNARY, to_api, uniform_data
from tests.test_adjoint_cpu import c5_population

pytestmark = pytest.mark.gpu

KW = dict(num_vars=3, max_tokens=127, max_depth=7, const_min=-2.0, const_max=2.0,
          unary_ops=GRAMMAR_UNARY, binary_ops=GRAMMAR_BINARY)


def test_evolve_step_matches_oracle(ctx):
    gp, op = GaParams.make(**KW), O.ga_params(**KW)
    cap = KW["max_tokens"] + KW["max_depth"] + 2
    n = 500
    rc, tok, val, ln = O.ga_init_population(op, n, cap, 41)
    assert rc == 0
    rng = np.random.default_rng(3)
    fit = rng.uniform(0, 10, n).astype(np.float32)
    fit[::17] = np.nan
    fit[5:40:7] = fit[4]  # ties
    pop = Population.from_strided(tok, val, ln, cap)
    for gen in (0, 3):
        t2, v2, l2, c2 = ctx.evolve_step(pop, fit, gp, SelectParams(7, 0.1, 0.05), 99, gen)
        assert c2 == cap
        rc, wt, wv, wl = O.ga_next_generation(op, 7, 0.1, 0.05, n, cap, tok, val, ln, fit, 99, gen)
        assert rc == 0 and np.array_equal(l2, wl)
        for i in range(n):
            s = slice(i * cap, i * cap + wl[i])
            assert np.array_equal(t2[s], wt[s]) and np.array_equal(v2[s].view(np.int32), wv[s].view(np.int32)), i
    bad = Population.from_exprs([([1, 0], [0.0, 1.0])])
    with pytest.raises(SrError) as e:
        ctx.evolve_step(bad, [1.0], gp, SelectParams(7, 0.1, 0.05), 1, 0)
    assert e.value.code == -3


def test_evolve_step_first_on_a_context_after_other_work():
    """sr_evolve_step as the first GA call of a context whose device memory
    was used before (evaluation, constant fitting): the GA status word starts
    cleared, so the step succeeds and matches the oracle."""
    from symbolic_regression_gpu_amd import EVAL_BATCHED, Context
    c = Context(0)
    try:
        nv, N = 4, 4096
        pop, slot = c5_population(64, 17, nv, min_len=20, max_tokens=120)
        X = uniform_data(nv, N, 6)
        c.set_data(X)
        api = to_api(pop)
        c.eval_mse(api, EVAL_BATCHED)
        fit = ConstantFit(c, api, slot)
        fit.adjoint_grad(fit.initial())
        torch.cuda.synchronize()
        gp, op = GaParams.make(**KW), O.ga_params(**KW)
        cap = KW["max_tokens"] + KW["max_depth"] + 2
        n = 300
        rc, tok, val, ln = O.ga_init_population(op, n, cap, 43)
        f = np.random.default_rng(4).uniform(0, 10, n).astype(np.float32)
        t2, v2, l2, _ = c.evolve_step(Population.from_strided(tok, val, ln, cap), f, gp,
                                      SelectParams(7, 0.1, 0.05), 5, 1)
        rc, wt, wv, wl = O.ga_next_generation(op, 7, 0.1, 0.05, n, cap, tok, val, ln, f, 5, 1)
        assert rc == 0 and np.array_equal(l2, wl)
        for i in range(n):
            s = slice(i * cap, i * cap + wl[i])
            assert np.array_equal(t2[s], wt[s]) and np.array_equal(v2[s].view(np.int32), wv[s].view(np.int32)), i
    finally:
        c.close()


def test_grad_consts_matches_fp64_bound_and_fit_path(ctx):
    nv, N = 4, 4096
    pop, slot = c5_population(64, 13, nv, min_len=20, max_tokens=120)
    X = uniform_data(nv, N, 5)
    X[nv] = np.sin(X[0]) * X[1] + X[2] - X[3]
    ctx.set_data(X)
    api = to_api(pop)
    g, mse = ctx.grad_consts(api, slot >= 0)
    assert np.all(g[slot < 0] == 0)
    # the constant-fitting path at the population's own constants: same bits
    fit = ConstantFit(ctx, api, slot)
    c0 = fit.initial()
    m2, g2 = fit.adjoint_grad(c0)
    torch.cuda.synchronize()
    m2, g2, c0 = m2.cpu().numpy(), g2.cpu().numpy(), c0.cpu().numpy()
    assert np.array_equal(mse, m2, equal_nan=True)
    covered = 0
    for e in range(pop.E):
        a, n = int(pop.off[e]), int(pop.len[e])
        sl = slot[a:a + n]
        ge = g[a:a + n]
        for k in range(8):
            hit = np.flatnonzero(sl == k)
            if len(hit):
                a_, b_ = ge[hit[0]], g2[e, k]
                assert a_ == b_ or (np.isnan(a_) and np.isnan(b_)), (e, k, a_, b_)
        tok, val = pop.expr(e)
        m64, g64, b64, cnt = O.mse_grad64_bound(tok, val, sl, X, nv, N, c0[e], max(N // 64, 1))
        if cnt == 0 or not np.isfinite(b64).all():
            continue
        covered += 1
        for k in range(8):
            hit = np.flatnonzero(sl == k)
            if len(hit):
                assert abs(ge[hit[0]] - g64[k]) <= 2 * b64[k], (e, k, ge[hit[0]], g64[k], b64[k])
    # the bound is finite only where no protected branch comes within rounding
    # at any of the N points (the same coverage the adjoint tests see)
    assert covered >= 8


def test_grad_consts_rejects_bad_masks(ctx):
    X = uniform_data(2, 1000, 1)
    ctx.set_data(X)
    pop = Population.from_exprs([([1, -1, 0], [0.0, 0.0, 2.0])])
    with pytest.raises(SrError) as e:
        ctx.grad_consts(pop, [0, 1, 0])  # a variable
    assert e.value.code == -1
    t = [1] * 9 + [0] * 10
    many = Population.from_exprs([(t, [0.0] * 9 + [float(k) for k in range(10)])])
    with pytest.raises(SrError) as e:
        ctx.grad_consts(many, [0] * 9 + [1] * 10)
    assert e.value.code == -7
    g, m = ctx.grad_consts(pop, [0, 0, 1])
    assert g[0] == 0 and g[1] == 0 and np.isfinite(g[2]) and np.isfinite(m[0])
