"""evogp population layout on the device vs the oracle, bit for bit, both
directions, and the GA population's import / export in that layout."""
import numpy as np
import pytest

from oracle import oracle as O
from symbolic_regression_gpu_amd.api import GaParams, SrError
from tests.helpers import GRAMMAR_BINARY, GRAMMAR_UNARY

pytestmark = pytest.mark.gpu

KW = dict(num_vars=3, max_tokens=127, max_depth=7, unary_ops=GRAMMAR_UNARY, binary_ops=GRAMMAR_BINARY)


def _cuda(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _host(t):
    return t.cpu().numpy().reshape(-1)


def _ga_pop(pop, seed):
    p = O.ga_params(**KW)
    cap = KW["max_tokens"] + KW["max_depth"] + 2
    rc, tok, val, ln = O.ga_init_population(p, pop, cap, seed)
    assert rc == 0
    return tok, val, ln, cap


def _same_evogp(a, b, ln, max_len):
    av, at, az = a
    bv, bt, bz = b
    for i, n in enumerate(ln):
        s = slice(i * max_len, i * max_len + max(int(n), 1))
        if not (0 < n <= max_len):
            assert az[i * max_len] == 0 and bz[i * max_len] == 0, i
            continue
        assert np.array_equal(av[s].view(np.int32), bv[s].view(np.int32)), i
        assert np.array_equal(at[s], bt[s]) and np.array_equal(az[s], bz[s]), i


def test_prefix_to_evogp_matches_oracle(ctx):
    tok, val, ln, cap = _ga_pop(500, 11)
    # a few broken streams: trailing tokens, a missing operand, an unknown token, too long
    tok[0 * cap:0 * cap + 4] = [1, 0, 0, 0]
    ln[0] = 4
    tok[1 * cap:1 * cap + 3] = [0, 0, 0]
    ln[1] = 3
    tok[2 * cap:2 * cap + 2] = [2, 0]
    ln[2] = 2
    tok[3 * cap:3 * cap + 2] = [28, 0]
    ln[3] = 2
    max_len = 64  # shorter than some individuals: those get size[0] = 0
    want = O.prefix_to_evogp(tok, val, ln, cap, max_len)
    got = ctx.prefix_to_evogp(_cuda(tok.reshape(-1, cap)), _cuda(val.reshape(-1, cap)), _cuda(ln), max_len)
    got = tuple(_host(t) for t in got)
    ok = np.array([0 < n <= max_len for n in ln]) & (want[2][::max_len] != 0)
    assert 0 < ok.sum() < len(ln) - 4
    assert list(want[2][:4 * max_len:max_len]) == [0, 0, 0, 0]
    _same_evogp(got, want, np.where(ok, ln, 0), max_len)


def test_evogp_to_prefix_matches_oracle_and_round_trips(ctx):
    tok, val, ln, cap = _ga_pop(400, 5)
    max_len = int(ln.max())
    value, type_, size = O.prefix_to_evogp(tok, val, ln, cap, max_len)
    type_[::max_len] |= 1 << 7  # output flag on every root: ignored
    size[7 * max_len] = -1  # skipped individuals
    size[8 * max_len] = max_len + 1
    want = O.evogp_to_prefix(value, type_, size, max_len, cap)
    got = ctx.evogp_to_prefix(_cuda(value.reshape(-1, max_len)), _cuda(type_.reshape(-1, max_len)),
                              _cuda(size.reshape(-1, max_len)), cap)
    gt, gv, gl = (_host(t) for t in got)
    assert np.array_equal(gl, want[2]) and gl[7] == 0 and gl[8] == 0
    for i, n in enumerate(gl):
        s = slice(i * cap, i * cap + n)
        assert np.array_equal(gt[s], want[0][s]) and np.array_equal(gv[s].view(np.int32), want[1][s].view(np.int32))
        if i not in (7, 8):
            assert np.array_equal(gt[s], tok[s])


def test_ga_import_export(ctx):
    gp = GaParams.make(**KW)
    ctx.ga_init(gp, 300, 77)
    tok, val, ln, cap = ctx.ga_population()
    max_len = int(ln.max())
    value, type_, size = ctx.ga_export_evogp(max_len)
    want = O.prefix_to_evogp(tok, val, ln, cap, max_len)
    _same_evogp(tuple(_host(t) for t in (value, type_, size)), want, ln, max_len)
    # import a different population (the oracle's), then read it back
    t2, v2, l2, _ = _ga_pop(256, 99)
    ev = O.prefix_to_evogp(t2, v2, l2, cap, int(l2.max()))
    ml = int(l2.max())
    ctx.ga_import_evogp(*(_cuda(a.reshape(-1, ml)) for a in ev))
    tok3, val3, ln3, cap3 = ctx.ga_population()
    assert cap3 == cap and np.array_equal(ln3, l2)
    for i, n in enumerate(l2):
        s = slice(i * cap, i * cap + n)
        assert np.array_equal(tok3[s], t2[s])
    # an individual with no tree is an error
    bad = [a.copy() for a in ev]
    bad[2][5 * ml] = 0
    with pytest.raises(SrError) as e:
        ctx.ga_import_evogp(*(_cuda(a.reshape(-1, ml)) for a in bad))
    assert e.value.code == -3
