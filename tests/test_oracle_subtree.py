"""The oracle's CPU subtree evaluator (FP64, cpu_subtree.cpp:83-501): the
secondary FP64 oracle for the cached path.

Pinned by
* tests/golden/subtree.npz: FP64 predictions of the reference's own
  eval_cpu_common_subtree_batch, compiled in place, on planted populations
  (bit for bit);
* the invariant every correct subtree cache satisfies: reusing a column
  computed from the same tokens in the same order changes no bit, so the
  predictions equal the direct FP64 evaluation (sro_predict64,
  cpu_simple_single.cpp:6-79) exactly, generation after generation;
* a brute-force count of the subtrees that join the stateful cache
  (size >= 2, frequency >= 5) and of the reuse on a repeated population.
The reference's stateful form hashes expression prefixes instead of subtrees
(detect.cpp:429-431); the fixture records its wrong predictions, which the
test shows differ from every correct evaluation.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import ALL_BINARY, collision_population, planted_population, random_population

HERE = os.path.dirname(os.path.abspath(__file__))
G = np.load(os.path.join(HERE, "golden", "subtree.npz"))
REF_UNARY = list(range(10, 28))


def _fixture_pop(g):
    tok, val, off, ln = (G[f"{k}{g}"] for k in ("tok", "val", "off", "len"))
    return O.Population([(tok[o:o + n], val[o:o + n]) for o, n in zip(off, ln)])


def _same(a, b):
    return np.array_equal(np.isnan(a), np.isnan(b)) and np.array_equal(a[~np.isnan(a)], b[~np.isnan(b)])


def _direct(pop, X32, nv, N):
    return np.stack([O.predict64(*pop.expr(e), X32, nv, N) for e in range(pop.E)])


@pytest.mark.parametrize("g", [0, 1])
def test_stateless_matches_reference_fixture(g):
    pop, X = _fixture_pop(g), G["X"]
    nv, N = X.shape[0] - 1, X.shape[1]
    for threads in (1, 5):
        pred, S = O.subtree_predict64(pop, X, nv, N, threads=threads)
        assert S > 0
        assert _same(pred, G[f"pred{g}"])
    # and a correct cache changes no bit of the direct FP64 evaluation
    assert _same(pred, _direct(pop, X.astype(np.float32), nv, N))


def test_reference_stateful_quirk_recorded():
    """detect_and_update_cache keys candidates by compute_hash(tokens[i],
    values[i], size): the expression's first `size` tokens, so different
    subtrees share a cache column and predictions go wrong.  The restatement
    keys each subtree by its own tokens and agrees with direct evaluation."""
    X = G["X"]
    nv, N = X.shape[0] - 1, X.shape[1]
    st = O.SubtreeState()
    wrong = 0
    for g in (0, 1):
        pop = _fixture_pop(g)
        direct = _direct(pop, X.astype(np.float32), nv, N)
        pred, _, _ = st.predict64(pop, X, nv, N)
        assert _same(pred, direct)
        wrong += sum(not _same(G[f"stateful{g}"][e], direct[e]) for e in range(pop.E))
    assert wrong > 0


def _brute_new(pops, min_size=2, min_freq=5):
    """Distinct subtrees (tokens + values) that join the cache, per call."""
    seen, out = set(), []
    for pop in pops:
        counts = {}
        for e in range(pop.E):
            t, v = pop.expr(e)
            for j in range(len(t)):
                s = O.lib().sro_subtree_size(np.ascontiguousarray(t, np.int32), j, len(t))
                if s >= min_size:
                    k = (tuple(t[j:j + s]), tuple(v[j:j + s].view(np.uint64)))
                    counts[k] = counts.get(k, 0) + 1
        new = {k for k, c in counts.items() if c >= min_freq and k not in seen}
        seen |= new
        out.append(len(new))
    return out


def test_stateful_over_generations_matches_direct_and_counts():
    nv, N = 3, 200
    rng = np.random.default_rng(3)
    X32 = rng.uniform(-2, 3, size=(nv + 1, N)).astype(np.float32)
    X = X32.astype(np.float64)
    pops = [planted_population(120, s, nv=nv, pool_size=12, unary=REF_UNARY) for s in (11, 12, 13)]
    want_new = _brute_new(pops)
    st = O.SubtreeState()
    total = 0
    for pop, wn in zip(pops, want_new):
        pred, new, hints = st.predict64(pop, X, nv, N, threads=4)
        assert new == wn
        total += new
        assert st.size == total
        assert _same(pred, _direct(pop, X32, nv, N))
        # every hint root spans a whole subtree; -2 only inside one
        for e in range(pop.E):
            t, _ = pop.expr(e)
            h = hints[pop.off[e]:pop.off[e] + len(t)]
            j = 0
            while j < len(t):
                s = O.lib().sro_subtree_size(np.ascontiguousarray(t, np.int32), j, len(t))
                if h[j] >= 0:
                    assert s >= 2 and np.all(h[j + 1:j + s] == -2)
                    j += s
                else:
                    assert h[j] == -1
                    j += 1
    # the same population again: no new entries, same predictions
    pred2, new2, _ = st.predict64(pops[-1], X, nv, N)
    assert new2 == 0 and _same(pred2, pred)
    # a different dataset size clears the state
    _, new3, _ = st.predict64(pops[0], X[:, :50], nv, 50)
    assert new3 == st.size == want_new[0]
    st.close()


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built here")
@pytest.mark.parametrize("seed", [21, 22])
def test_stateless_bit_exact_vs_compiled_reference(seed):
    nv, N = 4, 150
    rng = np.random.default_rng(seed)
    X = rng.uniform(-3, 5, size=(nv + 1, N))
    for pop in (planted_population(90, seed, nv=nv, unary=REF_UNARY),
                random_population(8, seed, nv=nv, max_tokens=40, max_depth=6, unary=REF_UNARY,
                                  binary=ALL_BINARY)):
        pred, _ = O.subtree_predict64(pop, X, nv, N)
        ref = np.empty((pop.E, N), np.float64)
        O.ref().ref_subtree_predict64(pop.tok, pop.val, pop.off, pop.len, pop.E, X.ravel(), nv, N, ref)
        assert _same(pred, ref)


def test_flat_hash_collision_keys_by_body():
    """The cross-generation cache (cl_gpu_eval.cu:384-397) keys slots by the
    flat hash alone in the reference; bodies are compared here."""
    pop = collision_population()
    a, b = pop.expr(0), pop.expr(1)
    assert O.hash_flat(a[0][2:5], a[1][2:5]) == O.hash_flat(b[0][2:5], b[1][2:5])
    ck = O.CacheKeys(16)
    ck.update(pop, 3, 5)
    keys = ck.keys
    # sin(x0*x0), sin(x2*x2), x0*x0, x2*x2: four slots, two of them with one key
    assert len(keys) == 4 and len(set(keys.tolist())) == 3
    h = ck.mark(pop, 3)
    assert h[pop.off[0] + 1] >= 0 and h[pop.off[1] + 1] >= 0 and h[pop.off[0] + 1] != h[pop.off[1] + 1]
    # with only the squares cached, each still gets its own slot
    ck2 = O.CacheKeys(16)
    sq = O.Population([pop.expr(e) for e in range(pop.E)] + [(np.array([3, -1, -1], np.int32),
                                                              np.array([0.0, 2.0, 2.0]))])
    ck2.update(O.Population([(t[2:5], v[2:5]) for t, v in (sq.expr(e) for e in range(sq.E))]), 3, 5)
    assert len(ck2.keys) == 2 and ck2.keys[0] == ck2.keys[1]
    h2 = ck2.mark(pop, 3)
    assert h2[pop.off[0] + 2] != h2[pop.off[1] + 2]
    ck.close()
    ck2.close()
    # the stateful FP64 evaluator keeps the two apart as well
    X = np.random.default_rng(0).uniform(1, 3, size=(4, 50)).astype(np.float32).astype(np.float64)
    st = O.SubtreeState()
    pred, new, _ = st.predict64(pop, X, 3, 50)
    assert new == 4  # x0*x0, x2*x2, sin(x0*x0), sin(x2*x2): six occurrences each
    assert _same(pred, _direct(pop, X.astype(np.float32), 3, 50))
    st.close()
