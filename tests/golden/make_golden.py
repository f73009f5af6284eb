"""Writes tests/golden/*.npz from the reference's own CPU code, compiled in
place by oracle/Makefile into oracle/_ref/libref.so (run in a container that
holds the reference tree):

    make -C oracle ref && python tests/golden/make_golden.py

Each fixture holds inputs and the reference's outputs on seeded data:
  ops.npz      eval_op_cpu (cl_cpu_eval.cpp:16-56) on special and random operands
  mse.npz      evaluate_cpu_mse (cl_cpu_eval.cpp:94-168) on a seeded population
  points.npz   evaluate_single_expr (cl_cpu_eval.cpp:59-90) per datapoint
  hashes.npz   compute_hash (detect.cpp:50-68) of seeded subtrees
  detect.npz   detect_common_subtrees (detect.cpp:120-387): sizes, counts,
               flat hashes of detected subtrees, hint map
  ga.npz       generate_random_expression / mutate_expression /
               crossover_expressions (utils.cpp:1156-1272) under set_seed(k)
  subtree.npz  eval_cpu_common_subtree_batch and eval_cpu_stateful_batch
               (cpu_subtree.cpp:135-501), FP64 predictions, on two planted
               populations sharing subtrees (python tests/golden/make_golden.py subtree)
  formats.npz  parse_input_info (utils.cpp:122-233) and load_data_file
               (utils.cpp:235-300) on the files under formats/: the
               reference's example digest and data files (data/examples/,
               copied as fixtures) and two synthetic files this script writes
               (odd spacing, exponents, signed zeros, CRLF rows)
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

OPS = list(range(1, 28))


def population(n, seed, nv, max_tokens, max_depth):
    p = O.ga_params(num_vars=nv, max_tokens=max_tokens, max_depth=max_depth,
                    const_min=-3.0, const_max=3.0)
    cap = max_tokens + max_depth + 2
    rc, tok, val, ln = O.ga_init_population(p, n, cap, seed)
    assert rc == 0
    return O.Population([(tok[i * cap:i * cap + ln[i]], val[i * cap:i * cap + ln[i]].astype(np.float64))
                         for i in range(n)])


def write_synthetic_formats():
    """Digest and data files in the reference's formats with awkward but valid
    number spellings; written once here and committed beside the outputs."""
    fdir = os.path.join(HERE, "formats")
    rng = np.random.default_rng(77)
    with open(os.path.join(fdir, "synth_digest.txt"), "w") as f:
        f.write("3\n")
        rows = [
            (2, 64, [1, -1, 10, -1], ["0", "1", "0", "0.0"]),
            (2, 64, [3, 0, 2, -1, 0], ["0", "-0.0", "0", "1", "1.5e-300"]),
            (2, 64, [8, 0, 0], ["0", "3.141592653589793238", "-2.5E+7"]),
        ]
        for nv, ndps, tok, val in rows:
            f.write(f"{nv} {ndps} {len(tok)}\n")
            f.write(" ".join(map(str, tok)) + "\n")
            f.write("  ".join(val) + "\n")
            f.write("data/synth_data.txt\n")
    with open(os.path.join(fdir, "synth_data.txt"), "w", newline="") as f:
        for j in range(64):
            x = rng.uniform(-3, 3, size=3)
            cells = [f"{x[0]:.17g}", f"{x[1]:.6e}", f"{x[2]:.3f}"]
            f.write("\t".join(cells) + ("\r\n" if j % 3 == 0 else "\n"))


def ref_digest(R, path, cap_e=64, cap_t=4096, stride=256):
    nv = np.zeros(cap_e, np.int32)
    nd = np.zeros(cap_e, np.int32)
    nt = np.zeros(cap_e, np.int32)
    tok = np.zeros(cap_t, np.int32)
    val = np.zeros(cap_t, np.float64)
    paths = C.create_string_buffer(cap_e * stride)
    shared = C.c_int()
    E = R.ref_parse_digest(path.encode(), cap_e, cap_t, stride, nv.ctypes.data, nd.ctypes.data,
                           nt.ctypes.data, tok.ctypes.data, val.ctypes.data, paths, C.byref(shared))
    T = int(nt[:E].sum())
    names = [paths.raw[e * stride:(e + 1) * stride].split(b"\0")[0].decode() for e in range(E)]
    return {"nv": nv[:E], "ndps": nd[:E], "ntok": nt[:E], "tok": tok[:T], "val": val[:T],
            "paths": np.array(names), "shared": shared.value}


def ref_data(R, path, nv, ndps):
    out = np.zeros((nv + 1) * ndps, np.float64)
    R.ref_load_data(path.encode(), nv, ndps, out.ctypes.data)
    return out.reshape(nv + 1, ndps)


def make_formats(R):
    fdir = os.path.join(HERE, "formats")
    write_synthetic_formats()
    out = {}
    for tag, name in (("sample", "sample_input.txt"), ("synth", "synth_digest.txt")):
        d = ref_digest(R, os.path.join(fdir, name))
        out.update({f"{tag}_{k}": v for k, v in d.items()})
    for tag, name, nv, ndps in (("s1", "sample_1", 1, 1000), ("s2", "sample_2", 3, 1000),
                                ("sd", "synth_data.txt", 2, 64), ("short", "synth_data.txt", 2, 70)):
        out[f"data_{tag}"] = ref_data(R, os.path.join(fdir, name), nv, ndps)
    np.savez(os.path.join(HERE, "formats.npz"), **out)


def make_subtree(R):
    from tests.helpers import planted_population
    nv, N = 3, 96
    rng = np.random.default_rng(77)
    X = rng.uniform(-3, 4, size=(nv + 1, N)).astype(np.float32).astype(np.float64)
    out = {"X": X}
    cache = R.ref_subtree_cache_create()
    try:
        for g, seed in enumerate((5, 6)):
            pop = planted_population(60, seed, nv=nv, pool_size=8, unary=list(range(10, 28)))
            pred = np.empty((pop.E, N), np.float64)
            R.ref_subtree_predict64(pop.tok, pop.val, pop.off, pop.len, pop.E, X.ravel(), nv, N, pred)
            spred = np.empty((pop.E, N), np.float64)
            n = R.ref_subtree_stateful_predict64(cache, pop.tok, pop.val, pop.off, pop.len, pop.E, X.ravel(), nv,
                                                 N, spred)
            out.update({f"tok{g}": pop.tok, f"val{g}": pop.val, f"off{g}": pop.off, f"len{g}": pop.len,
                        f"pred{g}": pred, f"stateful{g}": spred, f"cache_size{g}": np.int32(n)})
    finally:
        R.ref_subtree_cache_free(cache)
    np.savez(os.path.join(HERE, "subtree.npz"), **out)


def main():
    R = O.ref()
    if sys.argv[1:] == ["formats"]:
        make_formats(R)
        print("formats fixture written to", HERE)
        return
    if sys.argv[1:] == ["subtree"]:
        make_subtree(R)
        print("subtree fixture written to", HERE)
        return
    rng = np.random.default_rng(2024)
    # ---- op table -------------------------------------------------------
    special = np.array([0.0, -0.0, 1.0, -1.0, 0.5, -0.5, 2.0, 1e-9, -1e-9, 5e-10, -5e-10, 1e-10,
                        10.0, 10.5, 88.0, 89.0, -100.0, 1e30, -1e30, np.inf, -np.inf, np.nan],
                       np.float32)
    rnd = rng.uniform(-6, 6, size=12).astype(np.float32)
    vals = np.concatenate([special, rnd])
    a_list, b_list, op_list = [], [], []
    for op in OPS:
        for a in vals:
            for b in (vals if op < 10 else vals[:1]):
                op_list.append(op)
                a_list.append(a)
                b_list.append(b)
    ops = np.array(op_list, np.int32)
    a = np.array(a_list, np.float32)
    b = np.array(b_list, np.float32)
    out = np.array([R.ref_op32(int(o), float(x), float(y)) for o, x, y in zip(ops, a, b)], np.float32)
    np.savez(os.path.join(HERE, "ops.npz"), op=ops, a=a, b=b, out=out)

    # ---- MSE + per-point values ------------------------------------------
    pop = population(40, 77, nv=3, max_tokens=40, max_depth=6)
    N = 512
    X = rng.uniform(1, 5, size=(4, N)).astype(np.float32)
    X[3] = rng.normal(0, 2, size=N).astype(np.float32)
    mse = np.empty(pop.E, np.float32)
    R.ref_mse32(pop.tok, pop.val, pop.off, pop.len, pop.E, X.ravel(), 3, N, mse)
    np.savez(os.path.join(HERE, "mse.npz"), tok=pop.tok, val=pop.val, off=pop.off, len=pop.len,
             X=X, mse=mse)
    pts = np.empty((pop.E, 64), np.float32)
    for e in range(pop.E):
        t, v = pop.expr(e)
        for k in range(64):
            pts[e, k] = R.ref_eval_point32(np.ascontiguousarray(t), np.ascontiguousarray(v), len(t),
                                           np.ascontiguousarray(X[:3, k]), 3)
    np.savez(os.path.join(HERE, "points.npz"), tok=pop.tok, val=pop.val, off=pop.off, len=pop.len,
             X=X[:, :64].copy(), pred=pts)

    # ---- hashes ------------------------------------------------------------
    hp = population(30, 5, nv=4, max_tokens=30, max_depth=5)
    h = np.array([R.ref_hash_flat(*[np.ascontiguousarray(z) for z in hp.expr(e)], int(hp.len[e]))
                  for e in range(hp.E)], np.uint64)
    np.savez(os.path.join(HERE, "hashes.npz"), tok=hp.tok, val=hp.val, off=hp.off, len=hp.len,
             flat=h)

    # ---- detection: a population with planted shared subtrees ------------
    base = population(12, 9, nv=2, max_tokens=24, max_depth=5)
    exprs = [base.expr(e) for e in range(base.E)]
    shared = [([1, 3, -1, 0], [0, 0, 1, 2.5]), ([10, 2, -1, -1], [0, 0, 0, 1])]
    for k in range(24):
        t, v = shared[k % 2]
        bt, bv = exprs[k % base.E]
        exprs.append(([3] + list(t) + list(bt), [0.0] + list(v) + list(bv)))
    dpop = O.Population(exprs)
    maxs = 4096
    ss = np.zeros(maxs, np.int32)
    so = np.zeros(maxs, np.int32)
    sf = np.zeros(maxs, np.uint64)
    hints = np.zeros(int(dpop.len.sum()), np.int32)
    res = {}
    for (mn, mf) in ((3, 2), (3, 10), (2, 2)):
        n = R.ref_detect(dpop.tok, dpop.val, dpop.off, dpop.len, dpop.E, mn, mf, ss, so, sf, hints, maxs)
        res[f"n_{mn}_{mf}"] = np.int32(n)
        res[f"size_{mn}_{mf}"] = ss[:n].copy()
        res[f"occ_{mn}_{mf}"] = so[:n].copy()
        res[f"flat_{mn}_{mf}"] = sf[:n].copy()
        res[f"hints_{mn}_{mf}"] = hints.copy()
    np.savez(os.path.join(HERE, "detect.npz"), tok=dpop.tok, val=dpop.val, off=dpop.off,
             len=dpop.len, **res)

    # ---- GA operators under the reference's own mt19937 -------------------
    gp = O.ga_params(num_vars=3, max_tokens=48, max_depth=6, const_min=-2.0, const_max=2.0)
    parents = population(20, 31, nv=3, max_tokens=48, max_depth=6)
    # For each case: the reference's output, and the draw sequence (kind, lo,
    # hi, result) of its mt19937 + libstdc++ distributions, recorded through
    # the oracle run under the same engine (whose output is checked equal to
    # the reference's here).  A test replays the draws into the oracle.
    rows = []
    for seed in range(60):
        for op in range(3):
            a = parents.expr(seed % 20)
            bexp = parents.expr((seed * 7 + 3) % 20)
            at = np.ascontiguousarray(a[0], np.int32)
            av = np.ascontiguousarray(a[1], np.float32)
            bt = np.ascontiguousarray(bexp[0], np.int32)
            bv = np.ascontiguousarray(bexp[1], np.float32)
            t = np.zeros(4096, np.int32)
            v = np.zeros(4096, np.float32)
            n = R.ref_ga_op(op, 0, seed, C.byref(gp), at, av, len(at), bt, bv, len(bt), t, v)
            t2 = np.zeros(4096, np.int32)
            v2 = np.zeros(4096, np.float32)
            kind = np.zeros(8192, np.int32)
            lo = np.zeros(8192, np.float32)
            hi = np.zeros(8192, np.float32)
            res = np.zeros(8192, np.int32)
            nl = C.c_int()
            n2 = R.ref_ga_op_log(op, seed, C.byref(gp), at, av, len(at), bt, bv, len(bt), t2, v2,
                                 kind, lo, hi, res, 8192, C.byref(nl))
            assert n2 == n and np.array_equal(t[:n], t2[:n]) and np.array_equal(v[:n], v2[:n])
            assert nl.value <= 8192
            k = nl.value
            rows.append((seed, op, seed % 20, (seed * 7 + 3) % 20, t[:n].copy(), v[:n].copy(),
                         kind[:k].copy(), lo[:k].copy(), hi[:k].copy(), res[:k].copy()))
    np.savez(os.path.join(HERE, "ga.npz"),
             seed=np.array([r[0] for r in rows], np.int32), op=np.array([r[1] for r in rows], np.int32),
             a=np.array([r[2] for r in rows], np.int32), b=np.array([r[3] for r in rows], np.int32),
             out_len=np.array([len(r[4]) for r in rows], np.int32),
             out_tok=np.concatenate([r[4] for r in rows]), out_val=np.concatenate([r[5] for r in rows]),
             log_len=np.array([len(r[6]) for r in rows], np.int32),
             log_kind=np.concatenate([r[6] for r in rows]), log_lo=np.concatenate([r[7] for r in rows]),
             log_hi=np.concatenate([r[8] for r in rows]), log_res=np.concatenate([r[9] for r in rows]),
             p_tok=parents.tok, p_val=parents.val, p_off=parents.off, p_len=parents.len,
             params=np.array([3, 48, 6], np.int32), consts=np.array([-2.0, 2.0], np.float32))
    make_formats(R)
    make_subtree(R)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    O.build(ref=True)
    main()
