"""sr_eval_mse_noalloc (the allocation-free batched evaluation over the
library's static device scratch) against sr_eval_mse in SR_EVAL_BATCHED:
identical MSE bits and valid counts on two calls in a row, the second with a
smaller population (the scratch keeps the first call's contents), a malformed
expression, and the capacity error.  Runs on its own context."""
import ctypes as C

import numpy as np
import pytest

from symbolic_regression_gpu_amd import EVAL_BATCHED, Context, Population, SrError
from tests.helpers import GRAMMAR_BINARY, GRAMMAR_UNARY, random_population, to_api, uniform_data

pytestmark = pytest.mark.gpu


def _device(pop, torch):
    dev = torch.device("cuda", 0)
    tok = torch.from_numpy(np.ascontiguousarray(pop.tok, np.int32)).to(dev)
    val = torch.from_numpy(np.ascontiguousarray(pop.val, np.float32)).to(dev)
    off = torch.from_numpy(np.ascontiguousarray(pop.off, np.int32)).to(dev)
    ln = torch.from_numpy(np.ascontiguousarray(pop.len, np.int32)).to(dev)
    return tok, val, off, ln


def _noalloc(ctx, pop, torch):
    tok, val, off, ln = _device(pop, torch)
    mse = torch.full((pop.E,), 7.0, dtype=torch.float32, device="cuda")
    valid = torch.full((pop.E,), -5, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    ctx.eval_mse_noalloc(tok, val, off, ln, mse, valid)
    ctx.sync()
    return mse.cpu().numpy(), valid.cpu().numpy()


def test_noalloc_equals_batched_two_calls():
    torch = pytest.importorskip("torch")
    ctx = Context(0)
    try:
        N, nv = 100_003, 3
        X = uniform_data(nv, N, 31)
        X[nv] = np.cos(X[0]) * 2 + X[1]
        ctx.set_data(X)
        big = random_population(400, 61, nv=nv, max_tokens=120, max_depth=8, unary=GRAMMAR_UNARY,
                                binary=GRAMMAR_BINARY)
        small = random_population(37, 62, nv=nv, max_tokens=40, max_depth=6)
        for pop in (big, small):
            m_ref, v_ref, _ = ctx.eval_mse(to_api(pop), EVAL_BATCHED)
            m, v = _noalloc(ctx, pop, torch)
            assert np.array_equal(m.view(np.uint32), m_ref.view(np.uint32)), pop.E
            assert np.array_equal(v, v_ref), pop.E
            print(f"noalloc E={pop.E}: {np.isfinite(m).sum()} finite MSEs, bit-identical to sr_eval_mse")
    finally:
        ctx.close()


def test_noalloc_malformed_expression_and_capacity():
    torch = pytest.importorskip("torch")
    ctx = Context(0)
    try:
        N, nv = 5000, 2
        X = uniform_data(nv, N, 5)
        ctx.set_data(X)
        good = random_population(20, 71, nv=nv, max_tokens=30, max_depth=5)
        exprs = [good.expr(e) for e in range(good.E)]
        bad = (np.array([1, -1], np.int32), np.array([0.0, 1.0]))  # ADD with one operand
        from oracle import oracle as O
        pop = O.Population(exprs[:7] + [bad] + exprs[7:])
        m, v = _noalloc(ctx, pop, torch)
        m_ref, v_ref, _ = ctx.eval_mse(to_api(good), EVAL_BATCHED)
        keep = np.r_[0:7, 8:pop.E]
        assert np.array_equal(m[keep].view(np.uint32), m_ref.view(np.uint32))
        assert np.array_equal(v[keep], v_ref)
        assert np.isnan(m[7]) and v[7] == 0
        with pytest.raises(SrError):
            ctx.eval_mse(to_api(pop), EVAL_BATCHED)  # the allocating path rejects it
        # a layout beyond the static scratch is refused before anything runs
        tok, val, off, ln = _device(good, torch)
        out = torch.empty(good.E, dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        rc = ctx._L.sr_eval_mse_noalloc(ctx._h, good.E, 10_000_000, tok.data_ptr(), val.data_ptr(),
                                        off.data_ptr(), ln.data_ptr(), out.data_ptr(), None)
        assert rc == -5 and "static scratch" in ctx._L.sr_last_error(ctx._h).decode()
    finally:
        ctx.close()
