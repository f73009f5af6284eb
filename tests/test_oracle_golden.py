"""Pins the CPU oracle against fixtures written by the reference's own CPU code
(tests/golden/make_golden.py), plus closed forms the mathematics fixes.
Runs without a GPU and without the reference tree."""
import os

import numpy as np
import pytest

from oracle import oracle as O

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(G, name), allow_pickle=False)


def _pop(d, prefix=""):
    return O.Population([(d[prefix + "tok"][o:o + n], d[prefix + "val"][o:o + n])
                         for o, n in zip(d[prefix + "off"], d[prefix + "len"])])


def _bits_equal(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    return (np.isnan(a) & np.isnan(b)) | (a.view(np.int32) == b.view(np.int32))


def test_op_table_matches_reference():
    d = _load("ops.npz")
    got = np.array([O.op32(int(o), float(a), float(b)) for o, a, b in zip(d["op"], d["a"], d["b"])],
                   np.float32)
    ok = _bits_equal(got, d["out"])
    assert ok.all(), [(int(d["op"][i]), d["a"][i], d["b"][i], got[i], d["out"][i])
                      for i in np.flatnonzero(~ok)[:5]]


def test_mse_matches_reference():
    d = _load("mse.npz")
    pop = _pop(d)
    m, _ = O.mse32(pop, d["X"], 3, d["X"].shape[1], threads=3)
    assert _bits_equal(m, d["mse"]).all()


def test_points_match_reference():
    d = _load("points.npz")
    pop = _pop(d)
    X = d["X"]
    for e in range(pop.E):
        got = O.predict32(*pop.expr(e), X, 3, X.shape[1])
        assert _bits_equal(got, d["pred"][e]).all(), e


def test_flat_hashes_match_reference():
    d = _load("hashes.npz")
    pop = _pop(d)
    got = np.array([O.hash_flat(*pop.expr(e)) for e in range(pop.E)], np.uint64)
    assert np.array_equal(got, d["flat"])


@pytest.mark.parametrize("mn,mf", [(3, 2), (3, 10), (2, 2)])
def test_detection_matches_reference(mn, mf):
    d = _load("detect.npz")
    pop = _pop(d)
    r = O.detect(pop, mn, mf)
    n = int(d[f"n_{mn}_{mf}"])
    assert r["num_subs"] == n
    # The reference sorts with unstable std::sort, so compare as multisets
    # keyed by the subtree's flat hash.
    ref = sorted(zip(d[f"flat_{mn}_{mf}"].tolist(), d[f"size_{mn}_{mf}"].tolist(),
                     d[f"occ_{mn}_{mf}"].tolist()))
    got = sorted(zip(r["sub_flat"].tolist(), r["sub_size"].tolist(),
                     [len(o) for o in r["occurrences"]]))
    assert got == ref
    # Hint maps equal once subtree ids are replaced by the subtree's hash.
    ref_flat = d[f"flat_{mn}_{mf}"]

    def relabel(h, flats):
        return [int(flats[x]) if x >= 0 else x for x in h.tolist()]

    assert relabel(r["hints"], r["sub_flat"]) == relabel(d[f"hints_{mn}_{mf}"], ref_flat)
    assert 0.0 <= r["coverage"] <= 1.0


def test_ga_operators_replay_reference_draws():
    """Replays the reference's recorded draw sequence (mt19937 + libstdc++
    distributions, set_seed(k)) into the oracle's generate / mutate /
    crossover: every call must ask for the same (kind, lo, hi) in the same
    order, and the child must equal the reference's output bit for bit."""
    d = _load("ga.npz")
    nv, mt, md = (int(x) for x in d["params"])
    cmin, cmax = (float(x) for x in d["consts"])
    p = O.ga_params(num_vars=nv, max_tokens=mt, max_depth=md, const_min=cmin, const_max=cmax)
    par = O.Population([(d["p_tok"][o:o + n], d["p_val"][o:o + n])
                        for o, n in zip(d["p_off"], d["p_len"])])
    lo_off = np.concatenate([[0], np.cumsum(d["log_len"])])
    out_off = np.concatenate([[0], np.cumsum(d["out_len"])])
    for k in range(len(d["seed"])):
        kinds = d["log_kind"][lo_off[k]:lo_off[k + 1]]
        los = d["log_lo"][lo_off[k]:lo_off[k + 1]]
        his = d["log_hi"][lo_off[k]:lo_off[k + 1]]
        res = d["log_res"][lo_off[k]:lo_off[k + 1]]
        pos = [0]

        def take(kind, lo, hi):
            i = pos[0]
            assert i < len(kinds), "more draws than the reference made"
            assert kinds[i] == kind and np.float32(los[i]) == np.float32(lo) and \
                np.float32(his[i]) == np.float32(hi), (k, i)
            pos[0] += 1
            return res[i]

        uni = lambda lo, hi: float(np.array([take(0, lo, hi)], np.int32).view(np.float32)[0])  # noqa
        rint = lambda lo, hi: int(take(1, lo, hi))  # noqa
        a = par.expr(int(d["a"][k]))
        b = par.expr(int(d["b"][k]))
        t, v = O.ga_op_with_hooks(int(d["op"][k]), p, (a[0], a[1].astype(np.float32)),
                                  (b[0], b[1].astype(np.float32)), uni, rint)
        assert pos[0] == len(kinds), "fewer draws than the reference made"
        assert np.array_equal(t, d["out_tok"][out_off[k]:out_off[k + 1]]), k
        assert _bits_equal(v, d["out_val"][out_off[k]:out_off[k + 1]]).all(), k


# ---------------------------------------------------------------------------
# Closed forms and invariants (independent of any fixture)
# ---------------------------------------------------------------------------

def test_philox_known_answers():
    # Random123 published known-answer vectors for philox4x32-10.
    assert [hex(x) for x in O.philox([0, 0, 0, 0], [0, 0])] == \
        ["0x6627e8d5", "0xe169c58d", "0xbc57ac4c", "0x9b00dbd8"]
    assert [hex(x) for x in O.philox([0xffffffff] * 4, [0xffffffff] * 2)] == \
        ["0x408f276d", "0x41c83b0e", "0xa20bc7c6", "0x6d5451fd"]
    assert [hex(x) for x in O.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                                     [0xa4093822, 0x299f31d0])] == \
        ["0xd16cfe09", "0x94fdcceb", "0x5001e420", "0x24126ea1"]


def test_mse_closed_form_constant_expression():
    rng = np.random.default_rng(1)
    N = 1001
    X = rng.uniform(1, 5, size=(2, N)).astype(np.float32)
    c = 1.75
    pop = O.Population([([0], [c])])
    m, v = O.mse32(pop, X, 1, N)
    diff = (np.float32(c) - X[1]).astype(np.float32)
    want = np.sum((diff * diff).astype(np.float32).astype(np.float64)) / N
    assert v[0] == N and np.float32(want) == m[0]


def test_mse_excludes_nonfinite_and_nan_when_none():
    X = np.array([[0.0, 1.0, 2.0], [1.0, 1.0, 1.0]], np.float32)
    pop = O.Population([([18, -1], [0, 0]), ([4, 0, -1], [0, 1, 0]), ([4, -1, 0], [0, 0, 0.0])])
    m, v = O.mse32(pop, X, 1, 3)
    # INV(x): x=0 -> inf excluded; points 1, 2 give 1/1 - 1 = 0, 1/2 - 1 = -0.5
    assert v[0] == 2 and np.isclose(m[0], 0.125)
    # DIV(1, x): x = 0 -> NaN (cl_cpu_eval.cpp:24) excluded
    assert v[1] == 2
    # DIV(x, 0): all NaN -> MSE NaN
    assert v[2] == 0 and np.isnan(m[2])


def test_fp64_twin_agrees_on_exact_ops():
    # +,-,*,/ on small integers are exact in both precisions.
    for op, a, b, want in [(1, 3, 4, 7), (2, 3, 4, -1), (3, 3, 4, 12), (4, 3, 4, 0.75),
                           (6, 3, 4, 3), (7, 3, 4, 4), (24, -3, 0, 3), (25, 3, 0, -3)]:
        assert O.op32(op, a, b) == want
        assert O.op64(op, a, b) == want


def test_operand_order_left_then_right():
    X = np.zeros((1, 1), np.float32)
    for op, want in [(2, 6.0 - 2.0), (4, 6.0 / 2.0), (5, 36.0)]:
        pop = O.Population([([op, 0, 0], [0, 6.0, 2.0])])
        assert O.predict32(*pop.expr(0), X, 0, 1)[0] == np.float32(want)


def test_merkle_hash_composition():
    # hash(ADD(a, b)) = combine(combine(node(ADD), h(a)), h(b)) with the FNV
    # node hash and the boost-style combine of detect.cpp:97-118.
    FNV, P = 14695981039346656037, 1099511628211
    M = (1 << 64) - 1

    def node(tok, val):
        h = FNV
        h = ((h ^ (tok & M)) * P) & M
        h = ((h ^ int(np.array([val], np.float64).view(np.uint64)[0])) * P) & M
        return h

    def comb(h, c):
        return (h ^ ((c + 0x9E3779B97F4A7C15 + ((h << 6) & M) + (h >> 2)) & M)) & M

    hs, sz = O.merkle([1, -1, 0], [0.0, 1.0, 2.5])
    want = comb(comb(node(1, 0.0), node(-1, 1.0)), node(0, 2.5))
    assert int(hs[0]) == want and list(sz) == [3, 1, 1]
    assert O.hash_flat([-1], [1.0]) == node(-1, 1.0)


def test_detection_brute_force_small():
    # Against an O(n^2) brute-force count of equal subtrees.
    rng = np.random.default_rng(4)
    exprs = []
    pieces = [([1, -1, 0], [0, 0, 1.0]), ([10, -1], [0, 1.0]), ([3, 0, -1], [0, 2.0, 0])]
    for k in range(30):
        a = pieces[rng.integers(3)]
        b = pieces[rng.integers(3)]
        exprs.append(([2] + a[0] + b[0], [0.0] + a[1] + b[1]))
    pop = O.Population(exprs)
    r = O.detect(pop, 2, 2)
    occ = {}
    for e in range(pop.E):
        t, v = pop.expr(e)
        _, sz = O.merkle(t, v)
        for i in range(len(t)):
            if sz[i] >= 2:
                key = (tuple(t[i:i + sz[i]].tolist()), tuple(v[i:i + sz[i]].tolist()))
                occ[key] = occ.get(key, 0) + 1
    want = sorted(n for n in occ.values() if n >= 2)
    assert sorted(len(o) for o in r["occurrences"]) == want
