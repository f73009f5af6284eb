#!/usr/bin/env python3
"""Throughput of the MI355X evaluator on the evolution-loop workload:
population 1000, 1M datapoints, 3 variables ~ U[1,5], max depth 8, <= 255
tokens, tournament 7, crossover 0.9 / mutation 0.1, grammar {+,-,*,protected /}
x {sin, cos, clamped exp, protected log, sqrt|.|}, target a seeded depth-5
tree (the reference's evaluate -> fitness -> select -> vary step,
cl_sr_system.cpp:168-333).

One step = one generation: the population is evaluated by the batched
persistent interpreter over the generational subtree cache planned on the
device (hash, look up, evict, allocate, build and compile skeletons, two
interpreter launches, fused MSE), then ranked, selected and varied on the
device -- one captured hipGraph per generation, no host work.  `value` =
logical node-evals/s of the timed generations (sum_e len_e * N per generation
/ wall time); `ms_per_step` = ms per generation.  The same generations are
also run with the uncached batched evaluation replayed as one hipGraph per
generation and with the host-planned cache (their fitness vectors must be
identical), and the plain one-launch-per-expression interpreter is timed on
the initial and the final population.  Beside them:
  - single_expression: one tree (depth <= 6, <= 31 tokens, 3 variables) over
    100k datapoints: per-point values and MSE, plain and batched;
  - multi_expression: 100 trees of 5-60 tokens over 10 variables and 1M
    datapoints: plain vs batched, and the hipRTC module's compile latency
    (its device execution is switched off);
  - mutant_cache: a 36-45 token base tree and 100 mutants (one depth 1-3
    subtree replaced each), 1M datapoints: the reference-policy subtree cache
    (hit rate, FLOPs saved, node-evals/s) against the uncached evaluation;
  - constant_fit: 4096 expressions of 50-200 tokens over 3..8 of 8 variables,
    2-8 free constants, 256k datapoints: 20 Adam steps with the adjoint
    kernel vs central-difference gradient evaluations/s;
  - fp64_error (with the CPU leg): the GPU results against the oracle's FP64
    evaluator -- per-point ulps of the single expression, relative MSE errors
    of the batches on their first 65536 datapoints.

roofline: the cache interpreter kernel of the timed generations (HIP events
around each launch on the context stream): algorithmic HBM bytes per launch
(X|y, every column operand read and column value written) against the 8 TB/s
peak, with the flop basis (1 flop per operator token per datapoint, leaves
excluded, against the FP32 vector peak; no MFMA is issued) and the counted
VALU issue rate beside; roofline_batched is the uncached batched kernel on
the initial population, on the flop basis.  cpu_baseline: the oracle's restatement of
evaluate_cpu_mse (cl_cpu_eval.cpp:94-168) on a bounded sample of the initial
population, on this host's cores.

Usage: python bench.py [--gpus 1] [--steps 20] [--warmup 2]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

GRAMMAR_UNARY = [10, 11, 29, 22, 27]
GRAMMAR_BINARY = [1, 2, 3, 8]
METRIC = "node-evals/s + ms/generation (1000 exprs × 1M dps), plain vs batched+adjoint"
PEAK_FP32_TFLOPS = 157.3
DEVICE_CACHE_MIN_SIZE = 2
PEAK_HBM_GBS = 8000.0
TRAFFIC_FILE = os.path.join(HERE, "profiles", "traffic_k_vec_mse_evolution_1000x1M.json")
TRAFFIC_FILE_REF = os.path.join(HERE, "profiles", "traffic_k_vec_mseref_cached_evolution.json")
# VALU issue rate of both interpreter kernels from a counter run of this
# script (SQ_INSTS_VALU over the kernels' time against one wave64 VALU
# instruction per CU per cycle)
VALU_FILE = os.path.join(HERE, "profiles", "valu_issue_bench_default.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def operator_nodes(pop):
    return int(np.count_nonzero(pop.tokens > 0))


def to_population(tok, val, ln, cap):
    from symbolic_regression_gpu_amd import Population
    return Population.from_strided(tok, val, ln, cap)


def make_target(ctx, P, nv, N, depth, seed, min_len):
    """A seeded random tree of the grammar (device generator) and its values."""
    from symbolic_regression_gpu_amd.api import GaParams
    p = GaParams.make(num_vars=nv, max_tokens=255, max_depth=depth, const_min=-2.0, const_max=2.0,
                      unary_ops=GRAMMAR_UNARY, binary_ops=GRAMMAR_BINARY)
    for s in range(seed, seed + 64):
        ctx.ga_init(p, 64, s)
        tok, val, ln, cap = ctx.ga_population()
        for i in range(64):
            if ln[i] >= min_len:
                t = tok[i * cap:i * cap + ln[i]]
                v = val[i * cap:i * cap + ln[i]].astype(np.float64)
                return P.from_exprs([(t, v)])
    raise RuntimeError("no target tree of the requested length")


def evolution_setup(ctx, N, seed):
    from symbolic_regression_gpu_amd.api import GaParams
    nv = 3
    rng = np.random.default_rng(seed)
    X = rng.uniform(1.0, 5.0, size=(nv + 1, N)).astype(np.float32)
    ctx.set_data(X)
    from symbolic_regression_gpu_amd import Population
    target = make_target(ctx, Population, nv, N, 5, seed + 1, 15)
    X[nv] = ctx.eval_predictions(target)[0]
    X[nv] = np.nan_to_num(X[nv], nan=0.0, posinf=0.0, neginf=0.0)
    ctx.set_data(X)
    params = GaParams.make(num_vars=nv, max_tokens=255, max_depth=8, const_min=-2.0, const_max=2.0,
                           unary_ops=GRAMMAR_UNARY, binary_ops=GRAMMAR_BINARY)
    ctx.ga_init(params, 1000, seed + 2)
    tok, val, ln, cap = ctx.ga_population()
    return X, nv, (tok, val, ln, cap)


def time_evals(ctx, pop, mode, reps):
    """Mean logical node-evals/s and interpreter-kernel ms over reps calls
    (after one untimed call)."""
    ctx.eval_mse(pop, mode)
    wall, kms, launches, logical, executed = 0.0, 0.0, 0, 0, 0
    for _ in range(reps):
        t0 = time.perf_counter()
        _, _, st = ctx.eval_mse(pop, mode)
        wall += time.perf_counter() - t0
        kms += st["interp_kernel_ms"]
        launches += st["interp_launches"]
        logical += st["logical_node_evals"]
        executed += st["executed_node_evals"]
    return {"node_evals_per_s": logical / wall, "ms_per_eval": 1e3 * wall / reps,
            "interp_kernel_ms": kms / reps, "interp_launches": launches // reps,
            "executed_node_evals": executed // reps, "logical_node_evals": logical // reps}


def cpu_cores():
    """Threads for the CPU legs: the cores this process may run on, at most 16."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(X, nv, pop_tok, pop_val, pop_off, pop_len, target_s=12.0):
    """Oracle evaluate_cpu_mse restatement on a sample of the population: all
    expressions x the first n datapoints, n sized for ~target_s seconds."""
    from oracle import oracle as O
    cores = cpu_cores()
    N = X.shape[1]
    exprs = [(pop_tok[o:o + l], pop_val[o:o + l]) for o, l in zip(pop_off, pop_len)]
    pop = O.Population(exprs)
    tokens = int(pop.len.sum())

    def run(n):
        Xs = np.ascontiguousarray(X[:, :n])
        t0 = time.perf_counter()
        O.mse32(pop, Xs, nv, n, threads=cores)
        return time.perf_counter() - t0

    n = 2048
    dt = run(n)
    n = int(min(N, max(2048, n * target_s / max(dt, 1e-3))))
    dt = run(n)
    rate = tokens * n / dt
    return {"value": rate, "unit": "node-evals/s", "cores": cores, "kind": "port", "cpu_model": cpu_model(),
            "sample": f"all {len(exprs)} expressions ({tokens} tokens) x the first {n} of {N} "
                      f"datapoints, {dt:.1f} s, fp64-accumulated MSE as cl_cpu_eval.cpp:94-168"}


def constant_fit_bench(ctx, seed, adj_steps, fd_steps, E=4096, N=262144, nv=8):
    """Adjoint vs central-difference gradient evaluations/s (one grad-eval =
    dMSE/d(all free constants) of one expression over N points).  The
    expressions use 3..8 of the nv = 8 variables (one sixth each)."""
    import torch
    from symbolic_regression_gpu_amd import Population
    from symbolic_regression_gpu_amd.api import GaParams
    from symbolic_regression_gpu_amd.autograd import ConstantFit
    rng = np.random.default_rng(seed)
    X = rng.uniform(1.0, 5.0, size=(nv + 1, N)).astype(np.float32)
    ctx.set_data(X)
    target = make_target(ctx, Population, nv, N, 6, seed + 3, 20)
    y = ctx.eval_predictions(target)[0]
    X[nv] = np.nan_to_num(y, nan=0.0, posinf=0.0, neginf=0.0) + \
        rng.normal(0.0, 0.01, size=N).astype(np.float32)
    ctx.set_data(X)
    exprs, s = [], seed + 4
    for k, nv_e in enumerate(range(3, 9)):
        params = GaParams.make(num_vars=nv_e, max_tokens=200, max_depth=10, const_min=-2.0, const_max=2.0,
                               unary_ops=GRAMMAR_UNARY, binary_ops=GRAMMAR_BINARY)
        want = E * (k + 1) // 6
        while len(exprs) < want:
            ctx.ga_init(params, 8192, s)
            tok, val, ln, cap = ctx.ga_population()
            for i in range(8192):
                if ln[i] >= 50 and len(exprs) < want:
                    exprs.append((tok[i * cap:i * cap + ln[i]].copy(),
                                  val[i * cap:i * cap + ln[i]].astype(np.float64)))
            s += 1
    slot_rows, vals = [], []
    for t, v in exprs:
        sl = np.full(len(t), -1, np.int32)
        consts = np.flatnonzero(t == 0)
        k = min(len(consts), int(rng.integers(2, 9)))
        sl[consts[:k]] = np.arange(k, dtype=np.int32)
        v = v.copy()
        v[consts[:k]] = rng.normal(size=k).astype(np.float32)
        slot_rows.append(sl)
        vals.append(v)
    pop = Population.from_exprs([(t, v) for (t, _), v in zip(exprs, vals)])
    fit = ConstantFit(ctx, pop, np.concatenate(slot_rows))
    consts = fit.initial().requires_grad_()
    fwd_dpt, adj_dpt = ctx.param_autotune(consts.detach().data_ptr())
    opt = torch.optim.Adam([consts], lr=1e-2)

    def adam_step():
        opt.zero_grad()
        m = fit.mse(consts)
        m.nansum().backward()
        opt.step()
        return m

    m0 = adam_step().detach()  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(adj_steps):
        m = adam_step()
    torch.cuda.synchronize()
    t_adj = (time.perf_counter() - t0) / adj_steps
    fit.fd_grad(consts, 1e-3)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(fd_steps):
        fit.fd_grad(consts, 1e-3)
    torch.cuda.synchronize()
    t_fd = (time.perf_counter() - t0) / fd_steps
    m = m.detach()
    fin = torch.isfinite(m0) & torch.isfinite(m)
    return {"workload": f"{E} expressions of 50-200 tokens ({fit.tokens} tokens), 2-8 free "
                        f"constants, 3-8 of {nv} vars, {N} datapoints; Adam lr 1e-2",
            "adjoint_grad_evals_per_s": E / t_adj, "fd_grad_evals_per_s": E / t_fd,
            "adjoint_over_fd": t_fd / t_adj, "ms_per_adam_step": 1e3 * t_adj,
            "ms_per_fd_gradient": 1e3 * t_fd, "adam_steps_timed": adj_steps,
            "median_mse_first_last": [float(m0[fin].median()), float(m[fin].median())],
            "tuning": {"forward_dpt": fwd_dpt, "adjoint_dpt": adj_dpt}}


def random_exprs(ctx, nv, n, min_len, max_len, depth, seed, pool=4096):
    """n seeded random trees of the grammar from the device generator, with
    min_len <= length <= max_len (tokens, values as float64)."""
    from symbolic_regression_gpu_amd.api import GaParams
    p = GaParams.make(num_vars=nv, max_tokens=max_len, max_depth=depth, const_min=-2.0, const_max=2.0,
                      unary_ops=GRAMMAR_UNARY, binary_ops=GRAMMAR_BINARY)
    out, s = [], seed
    while len(out) < n:
        ctx.ga_init(p, pool, s)
        tok, val, ln, cap = ctx.ga_population()
        for i in range(pool):
            if min_len <= ln[i] <= max_len and len(out) < n:
                out.append((tok[i * cap:i * cap + ln[i]].copy(), val[i * cap:i * cap + ln[i]].astype(np.float64)))
        s += 1
    return out


def subtree_sizes_depths(tok):
    """Size (tokens) and depth of the subtree rooted at every prefix position."""
    L = len(tok)
    size = np.zeros(L, np.int64)
    depth = np.zeros(L, np.int64)
    for i in range(L - 1, -1, -1):
        t = int(tok[i])
        s, d, pos = 1, 1, i + 1
        for _ in range(0 if t <= 0 else (2 if t < 10 else 1)):
            s += size[pos]
            d = max(d, 1 + depth[pos])
            pos += size[pos]
        size[i], depth[i] = s, d
    return size, depth


def timed(fn, reps):
    """Mean wall ms of fn() over reps calls after one untimed call, and its last result."""
    fn()
    t0 = time.perf_counter()
    for _ in range(reps):
        r = fn()
    return 1e3 * (time.perf_counter() - t0) / reps, r


def single_expression_bench(ctx, seed, reps=50):
    """One seeded tree (depth <= 6, <= 31 tokens) over 3 variables ~ U[1,5] and
    100k datapoints, target a second tree of the same grammar: per-point
    values and MSE on the GPU (plain and batched interpreters)."""
    from symbolic_regression_gpu_amd import EVAL_BATCHED, EVAL_PLAIN, Population
    nv, N = 3, 100_000
    rng = np.random.default_rng(seed)
    X = rng.uniform(1.0, 5.0, size=(nv + 1, N)).astype(np.float32)
    ctx.set_data(X)
    model, target = random_exprs(ctx, nv, 2, 15, 31, 6, seed)
    X[nv] = np.nan_to_num(ctx.eval_predictions(Population.from_exprs([target]))[0], nan=0.0, posinf=0.0,
                          neginf=0.0)
    ctx.set_data(X)
    pop = Population.from_exprs([model])
    ms_pred, pred = timed(lambda: ctx.eval_predictions(pop), reps)
    res = {"workload": f"one tree of {len(model[0])} tokens (depth <= 6), 3 vars, {N} datapoints",
           "predictions_ms": ms_pred}
    for name, mode in (("plain", EVAL_PLAIN), ("batched", EVAL_BATCHED)):
        ms, (m, v, st) = timed(lambda: ctx.eval_mse(pop, mode), reps)
        res[name] = {"ms_per_eval": ms, "node_evals_per_s": len(model[0]) * N / (ms * 1e-3),
                     "interp_kernel_ms": st["interp_kernel_ms"], "mse": float(m[0]), "valid": int(v[0])}
    return res, (X, nv, [model], pred[0])


def multi_expression_bench(ctx, seed, reps=5):
    """100 seeded trees of 5-60 tokens over 10 variables ~ U[1,5], 1M
    datapoints: the plain one-launch-per-expression interpreter, the batched
    interpreter, and the hipRTC backend's compile latency for the same batch."""
    from symbolic_regression_gpu_amd import EVAL_BATCHED, EVAL_PLAIN, Population
    from symbolic_regression_gpu_amd import api
    nv, N = 10, 1 << 20
    rng = np.random.default_rng(seed)
    X = rng.uniform(1.0, 5.0, size=(nv + 1, N)).astype(np.float32)
    exprs = random_exprs(ctx, nv, 101, 5, 60, 7, seed)
    ctx.set_data(X)
    X[nv] = np.nan_to_num(ctx.eval_predictions(Population.from_exprs([exprs[100]]))[0], nan=0.0,
                          posinf=0.0, neginf=0.0)
    ctx.set_data(X)
    exprs = exprs[:100]
    pop = Population.from_exprs(exprs)
    res = {"workload": f"100 trees of 5-60 tokens ({pop.total_tokens} tokens), 10 vars, {N} datapoints"}
    for name, mode in (("plain", EVAL_PLAIN), ("batched", EVAL_BATCHED)):
        res[name] = time_evals(ctx, pop, mode, reps)
    res["batched_over_plain"] = res["batched"]["node_evals_per_s"] / res["plain"]["node_evals_per_s"]
    t0 = time.perf_counter()
    compile_ms, code_bytes = api.jit_compile(pop, nv, 4)
    res["jit"] = {"compile_ms": compile_ms, "compile_wall_ms": 1e3 * (time.perf_counter() - t0),
                  "code_bytes": code_bytes,
                  "eval_ms": None,
                  "note": "hipRTC module compiled on the host; its device execution is switched off "
                          "(the first GPU run stopped making progress; cause not identified)"}
    return res, (X, nv, exprs)


def mutant_cache_bench(ctx, seed, reps=5):
    """One seeded base tree (36-45 tokens, 5 variables) and 100 mutants, each
    with one subtree of depth 1-3 replaced by a fresh random subtree of depth
    1-3; 1M datapoints ~ U[1,5]: the reference-policy subtree cache (detect,
    cache shared subtrees, skeletons) against the uncached batched evaluation."""
    from symbolic_regression_gpu_amd import EVAL_BATCHED, EVAL_CACHED, Population
    nv, N = 5, 1 << 20
    rng = np.random.default_rng(seed)
    X = rng.uniform(1.0, 5.0, size=(nv + 1, N)).astype(np.float32)
    base, target = random_exprs(ctx, nv, 2, 36, 45, 7, seed)
    fresh = random_exprs(ctx, nv, 100, 2, 15, 3, seed + 1)
    bt, bv = base
    size, depth = subtree_sizes_depths(bt)
    cand = np.flatnonzero((depth >= 1) & (depth <= 3) & (np.arange(len(bt)) > 0))
    exprs = [base]
    for k in range(100):
        i = int(rng.choice(cand))
        ft, fv = fresh[k]
        exprs.append((np.concatenate([bt[:i], ft, bt[i + size[i]:]]),
                      np.concatenate([bv[:i], fv, bv[i + size[i]:]])))
    ctx.set_data(X)
    X[nv] = np.nan_to_num(ctx.eval_predictions(Population.from_exprs([target]))[0], nan=0.0, posinf=0.0,
                          neginf=0.0)
    ctx.set_data(X)
    pop = Population.from_exprs(exprs)
    logical = pop.total_tokens * N
    unc = time_evals(ctx, pop, EVAL_BATCHED, reps)
    m_b, _, _ = ctx.eval_mse(pop, EVAL_BATCHED)
    ctx.cache_config(2048, 3, 2)
    ctx.cache_clear()
    t0 = time.perf_counter()
    m_c, _, st1 = ctx.eval_mse(pop, EVAL_CACHED)
    first_ms = 1e3 * (time.perf_counter() - t0)
    steady = time_evals(ctx, pop, EVAL_CACHED, reps)
    ctx.cache_clear()
    return {"workload": f"base tree of {len(bt)} tokens + 100 mutants ({pop.total_tokens} tokens), 5 vars, "
                        f"{N} datapoints; reference-policy cache, min size 3, min frequency 2",
            "uncached_batched": unc,
            "cached_first_call": {"ms": first_ms, "subtrees_detected": st1["num_subtrees"],
                                  "subtrees_cached": st1["num_cached"], "coverage": st1["coverage"],
                                  "executed_over_logical": st1["executed_node_evals"] / logical},
            "cached_steady": steady,
            "cache_hit_rate": st1["coverage"],
            "flops_saved": 1.0 - steady["executed_node_evals"] / logical,
            "cached_over_uncached": steady["node_evals_per_s"] / unc["node_evals_per_s"],
            "mse_identical_to_uncached": bool(np.array_equal(m_b.view(np.uint32), m_c.view(np.uint32)))}, \
        (X, nv, exprs)


def fp64_accuracy(ctx, cores, c1, c2, c3, evo, n_sample=1 << 16):
    """Error of the GPU results against the oracle's FP64 evaluator, the
    checker (sro_predict64 / sro_mse64 restating cl_cpu_eval.cpp in double):
    per-point values of the single expression (all its datapoints) in fp32
    ulps of the FP64 value, and relative MSE errors of the batches on their
    first n_sample datapoints (the GPU evaluates the same sample), with the
    share within 1e-5 of the FP64 MSE and of the oracle's fp32 MSE (the
    reference CPU evaluator's arithmetic; an ill-conditioned expression is
    off FP64 in both)."""
    from oracle import oracle as O
    from symbolic_regression_gpu_amd import EVAL_BATCHED, Population
    out = {}
    X, nv, exprs, pred = c1
    t, v = exprs[0]
    N = X.shape[1]
    v64 = O.predict64(t, v, X, nv, N)
    c32 = O.predict32(t, v, X, nv, N)
    g = np.asarray(pred, np.float64)
    fin = np.isfinite(v64) & np.isfinite(g)
    with np.errstate(invalid="ignore", over="ignore"):
        ulps = np.abs(g[fin] - v64[fin]) / np.spacing(np.abs(v64[fin]).astype(np.float32)).astype(np.float64)
        near32 = np.abs(g - c32) <= 4 * np.spacing(np.abs(c32).astype(np.float32)).astype(np.float64)
    same_class = (np.isnan(g) == np.isnan(c32)) & (np.isinf(g) == np.isinf(c32))
    out["single_expression_points"] = {
        "points": int(N), "max_ulp_vs_fp64": float(ulps.max()) if ulps.size else 0.0,
        "median_ulp_vs_fp64": float(np.median(ulps)) if ulps.size else 0.0,
        "share_within_4ulp_of_oracle_fp32": float(np.mean(same_class & (near32 | ~np.isfinite(c32))))}
    for name, (Xb, nvb, eb) in (("multi_expression", c2), ("mutant_cache", c3), ("evolution_final_population", evo)):
        n = min(n_sample, Xb.shape[1])
        Xs = np.ascontiguousarray(Xb[:, :n])
        ctx.set_data(Xs)
        m_gpu, _, _ = ctx.eval_mse(Population.from_exprs(eb), EVAL_BATCHED)
        opop = O.Population(eb)
        m64 = O.mse64(opop, Xs, nvb, n, threads=cores)
        m32, _ = O.mse32(opop, Xs, nvb, n, threads=cores)
        m_gpu = np.asarray(m_gpu, np.float64)
        m32 = np.asarray(m32, np.float64)
        both = np.isfinite(m_gpu) & np.isfinite(m64) & (m64 > 0)
        rel = np.abs(m_gpu[both] - m64[both]) / m64[both]
        b32 = np.isfinite(m_gpu) & np.isfinite(m32) & (m32 > 0)
        rel32 = np.abs(m_gpu[b32] - m32[b32]) / m32[b32]
        out[name] = {"exprs": len(eb), "datapoints": n, "finite": int(both.sum()),
                     "max_rel_mse_err": float(rel.max()) if rel.size else 0.0,
                     "median_rel_mse_err": float(np.median(rel)) if rel.size else 0.0,
                     "share_within_1e-5": float(np.mean(rel <= 1e-5)) if rel.size else 1.0,
                     "share_within_1e-5_of_oracle_fp32": float(np.mean(rel32 <= 1e-5)) if rel32.size else 1.0,
                     "nan_class_agrees": float(np.mean(np.isnan(m_gpu) == np.isnan(m64)))}
    return out


def copy_calibration(mib):
    """A device copy of known size, for scaling FETCH_SIZE / WRITE_SIZE."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, device="cuda", dtype=torch.float32)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4}))


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--N", type=int, default=1 << 20)
    ap.add_argument("--reps", type=int, default=5, help="timed batched/plain evaluations")
    ap.add_argument("--adj-steps", type=int, default=20)
    ap.add_argument("--fd-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fit", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MiB")
    ap.add_argument("--roofline-only", action="store_true",
                    help="only the timed batched evaluations of the initial population (for counter runs)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the evaluator would shard datapoints with one all-reduce of per-expression "
              "partial sums per generation).", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return

    import torch
    from symbolic_regression_gpu_amd import EVAL_BATCHED, EVAL_CACHED, EVAL_PLAIN, Context
    from symbolic_regression_gpu_amd.api import (CACHE_DEVICE, CACHE_GENERATIONAL_SHARED, CACHE_REFERENCE,
                                                 SelectParams)
    torch.cuda.init()
    ctx = Context(0)
    N = args.N
    K, W = args.steps, args.warmup
    log(f"[bench] evolution workload: pop 1000, N={N}")
    X, nv, (tok, val, ln, cap) = evolution_setup(ctx, N, args.seed)
    pop0 = to_population(tok, val, ln, cap)
    flops0 = operator_nodes(pop0) * N
    log("[bench] autotuning the batched launch shape")
    tuned, tune_ms = ctx.autotune(pop0, max_sample=256)
    xy_bytes = (nv + 1) * 4 * N

    log("[bench] batched evaluation of the initial population")
    batched0 = time_evals(ctx, pop0, EVAL_BATCHED, args.reps)
    if args.roofline_only:
        print(json.dumps({"kernel_ms": batched0["interp_kernel_ms"], "flops_per_launch": flops0,
                          "launches": args.reps + 1, "tuning": [tuned.dpt, tuned.chunk, tuned.n_queues,
                                                                tuned.blocks_per_cu]}), flush=True)
        return
    log("[bench] plain one-launch-per-expression evaluation of the initial population")
    plain0 = time_evals(ctx, pop0, EVAL_PLAIN, max(2, args.reps // 2))
    log("[bench] reference-policy subtree cache on the initial population")
    ctx.cache_config(2048, 3, 10)
    ctx.cache_clear()
    _, _, st_c1 = ctx.eval_mse(pop0, EVAL_CACHED)
    ref_cache = time_evals(ctx, pop0, EVAL_CACHED, 2)
    ref_cache["first_call_new_subtrees"] = st_c1["num_new_subtrees"]
    ref_cache["coverage"] = st_c1["coverage"]

    sel = SelectParams(tournament=7, p_mutation=0.1, elite_frac=0.05)
    # (A) device-resident generations, batched evaluation, one hipGraph each
    ctx.ga_set_population(tok, val, ln, cap)
    log(f"[bench] batched generations: {W} warm-up + {K} timed (hipGraph)")
    ctx.ga_run(sel, EVAL_BATCHED, W, 0, args.seed, True)
    ctx.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist_a, ms_dev, st_a = ctx.ga_run(sel, EVAL_BATCHED, K, W, args.seed, True)
    ctx.sync()
    torch.cuda.synchronize()
    wall_a = time.perf_counter() - t0
    batched_gen = {"node_evals_per_s": st_a["logical_node_evals"] / wall_a,
                   "ms_per_generation": 1e3 * wall_a / K,
                   "device_ms_per_generation": ms_dev / K}

    # (B) the timed steps: the same generations (same seed, same initial
    # population) with the generational subtree cache planned on the device,
    # one captured graph per generation; its MSEs equal the batched ones bit
    # for bit, so selection and offspring are identical.
    # subtrees of >= 2 tokens get columns (a unary op over a leaf included):
    # measured the fastest minimum size on this workload (2..7)
    ctx.cache_policy(CACHE_DEVICE, 0, DEVICE_CACHE_MIN_SIZE)
    ctx.ga_set_population(tok, val, ln, cap)
    log(f"[bench] cached generations (device-planned cache, hipGraph): {W} warm-up + {K} timed")
    ctx.ga_run(sel, EVAL_CACHED, W, 0, args.seed, True)
    # skeleton kernel tuned on the last warm-up generation's plan
    _, cache_tune_ms = ctx.autotune_cache()
    tuned = ctx.tuning
    ctx.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist_b, ms_dev_b, st_b = ctx.ga_run(sel, EVAL_CACHED, K, W, args.seed, True)
    ctx.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    value = st_b["logical_node_evals"] / wall
    identical = bool(np.array_equal(hist_a, hist_b, equal_nan=True))
    best = float(np.nanmin(np.where(np.isfinite(hist_b[-1]), hist_b[-1], np.nan)))
    popK = to_population(*ctx.ga_population())
    # the same generations again without graphs: HIP events around each of
    # the two interpreter launches (the roofline's kernel time)
    log("[bench] the same cached generations, launches timed with events")
    ctx.cache_policy(CACHE_DEVICE, 0, DEVICE_CACHE_MIN_SIZE)
    ctx.ga_set_population(tok, val, ln, cap)
    ctx.ga_run(sel, EVAL_CACHED, W, 0, args.seed, False)
    hist_t, _, st_t = ctx.ga_run(sel, EVAL_CACHED, K, W, args.seed, False)
    assert np.array_equal(hist_t, hist_b, equal_nan=True)
    cached_gen = {"node_evals_per_s": value, "ms_per_generation": 1e3 * wall / K,
                  "device_ms_per_generation": ms_dev_b / K,
                  "executed_over_logical": st_b["executed_node_evals"] / max(st_b["logical_node_evals"], 1),
                  "interp_kernel_ms_per_generation": st_t["interp_kernel_ms"] / K,
                  "launches_per_generation": st_t["interp_launches"] / K,
                  "columns_written_per_generation": st_b["column_writes"] / N / K,
                  "columns_read_per_generation": st_b["column_reads"] / N / K,
                  "columns_live": st_b["num_cached"],
                  "host_work_per_generation": "none (planned on the device, one hipGraph per generation)",
                  "fitness_identical_to_batched": identical}

    # (B2) for comparison: the host-planned generational cache (SHARED policy)
    ctx.cache_policy(CACHE_GENERATIONAL_SHARED, 0, 3)
    ctx.ga_set_population(tok, val, ln, cap)
    log(f"[bench] cached generations (host-planned cache): {W} warm-up + {K} timed")
    ctx.ga_run(sel, EVAL_CACHED, W, 0, args.seed, False)
    ctx.sync()
    t0 = time.perf_counter()
    hist_h, _, st_h = ctx.ga_run(sel, EVAL_CACHED, K, W, args.seed, False)
    ctx.sync()
    wall_h = time.perf_counter() - t0
    host_cached = {"node_evals_per_s": st_h["logical_node_evals"] / wall_h, "ms_per_generation": 1e3 * wall_h / K,
                   "host_plan_ms_per_generation": st_h["detect_time_ms"] / K,
                   "executed_over_logical": st_h["executed_node_evals"] / max(st_h["logical_node_evals"], 1),
                   "fitness_identical_to_batched": bool(np.array_equal(hist_a, hist_h, equal_nan=True))}
    ctx.cache_policy(CACHE_REFERENCE, 0, 4)

    # (C) plain and batched on the last population of the timed run
    log("[bench] plain and batched evaluation of the final population")
    plainK = time_evals(ctx, popK, EVAL_PLAIN, 2)
    batchedK = time_evals(ctx, popK, EVAL_BATCHED, 3)
    evo = (X, nv, [popK.expr(e) for e in range(popK.num_exprs)])

    log("[bench] single expression, 100k datapoints")
    c1, c1_data = single_expression_bench(ctx, args.seed + 20)
    log("[bench] 100 expressions over 10 variables, 1M datapoints: plain, batched, hipRTC compile")
    c2, c2_data = multi_expression_bench(ctx, args.seed + 30)
    log("[bench] base tree + 100 mutants, 1M datapoints: subtree cache vs uncached")
    c3, c3_data = mutant_cache_bench(ctx, args.seed + 40)

    fit = None
    if not args.no_fit:
        log("[bench] constant fitting: adjoint vs finite differences")
        fit = constant_fit_bench(ctx, args.seed + 10, args.adj_steps, args.fd_steps)

    cpu = acc = None
    if not args.no_cpu:
        log("[bench] CPU baseline")
        cpu = cpu_baseline(X, nv, pop0.tokens, pop0.values, pop0.offsets, pop0.lengths)
        log("[bench] error against the FP64 evaluator")
        acc = fp64_accuracy(ctx, cpu["cores"], c1_data, c2_data, c3_data, evo)

    # roofline of the dominant kernel of the timed steps: the skeleton
    # interpreter (MODE_MSE_REF), over every launch of the timed
    # generations (HIP events around each launch on the context stream)
    launches = max(st_t["interp_launches"], 1)
    k_ms = st_t["interp_kernel_ms"] / launches
    flops = st_t["executed_op_evals"] / launches
    achieved = flops / (k_ms * 1e-3) / 1e12
    hbm_alg = (st_t["column_reads"] + st_t["column_writes"]) * 4 / launches + xy_bytes
    def profiled_traffic(path, kernel):
        """HBM bytes per launch of a FETCH_SIZE + WRITE_SIZE counter run, if it covers this kernel."""
        if not os.path.exists(path):
            return None
        with open(path) as f:
            tj = json.load(f)
        return tj.get("hbm_bytes_per_launch") if tj.get("kernel", "").startswith(kernel.split("<")[0]) else None

    traffic = None
    b_ms = batched0["interp_kernel_ms"]
    b_ach = flops0 / (b_ms * 1e-3) / 1e12
    def xc_of(dpt):
        """Variable-register class of the batched kernel variant (<= 4 variables)."""
        return 3 if nv <= 3 and dpt == 3 else 4

    def kernel_name(kind, dpt, mode):
        """Name of the interpreter kernel a launch plan runs (as the kernel trace shows it)."""
        if kind == 1:
            return f"k_vec<{mode}>"
        return f"k_batched<{dpt},9,{xc_of(dpt)},{mode}>"

    cache_dpt = tuned.cache_dpt if tuned.cache_dpt > 0 else 3
    ref_kernel = kernel_name(tuned.cache_kernel, cache_dpt, "MODE_MSE_REF")
    mse_kernel = kernel_name(tuned.kernel, tuned.dpt, "MODE_MSE")
    valu_ref = valu_mse = None
    if os.path.exists(VALU_FILE):
        with open(VALU_FILE) as f:
            vj = json.load(f)
        for k, v in vj.items():
            if not isinstance(v, dict):
                continue
            issue = {"achieved": v["valu_instr_per_s"], "peak": v["valu_issue_peak_per_s"],
                     "unit": "VALU wave64 instructions/s", "frac": v["valu_issue_frac"],
                     # the scalar unit of a CU issues at most one instruction per cycle too
                     "salu_frac": v.get("salu_issue_frac"),
                     "source": "profiles/valu_issue_bench_default.json (counter run)"}
            # a counter run describes the kernel it names, nothing else
            kk = k.replace(" ", "")
            if kk.startswith(ref_kernel.split("<")[0]) and "MODE_MSE_REF" in kk:
                valu_ref = issue
            elif kk.startswith(mse_kernel.split("<")[0]) and "MODE_MSE" in kk and "MODE_MSE_REF" not in kk:
                valu_mse = issue
    traffic = profiled_traffic(TRAFFIC_FILE_REF, ref_kernel)
    b_traffic = profiled_traffic(TRAFFIC_FILE, mse_kernel)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "node-evals/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": 1e3 * wall / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded Philox GP population and target tree, X ~ U[1,5] fp32",
        "config": {
            "workload": "evolution loop: population 1000, 1M datapoints, 3 vars, max depth 8, "
                        "<=255 tokens, tournament 7, crossover 0.9 / mutation 0.1, elites 5%, "
                        "target depth-5 tree; one step = one generation: batched persistent "
                        "evaluation over the generational subtree cache planned on the device, "
                        "fused MSE, device select/vary, one hipGraph per generation",
            "N": N, "population": 1000, "tokens_initial": int(pop0.tokens.size),
            "tokens_final": int(popK.tokens.size),
            "node_evals_per_generation": st_b["logical_node_evals"] // max(K, 1),
            "best_mse_last_generation": best,
            "tuning": {"kernel": mse_kernel, "cache_kernel": ref_kernel, "chunk": tuned.chunk,
                       "n_queues": tuned.n_queues, "blocks_per_cu": tuned.blocks_per_cu,
                       "cache_kernel_ms_on_tuning_plan": cache_tune_ms},
        },
        # algorithmic HBM bytes per launch (X|y once, every column operand
        # read and every column value written) against the HBM peak; the
        # kernel issues no MFMA, and its limiter is VALU issue / latency:
        # the flop basis and the counted VALU issue rate are given beside
        "roofline": {
            "bound": "hbm",
            "achieved": hbm_alg / (k_ms * 1e-3) / 1e9,
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": hbm_alg / (k_ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
            "traffic": traffic,
            "kernel": f"{ref_kernel} (cache columns and skeletons), average over the timed generations' launches",
            "kernel_ms": k_ms,
            "launches": launches,
            "hbm_algorithmic_bytes_per_launch": hbm_alg,
            "flop_basis": {"achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                           "frac": achieved / PEAK_FP32_TFLOPS, "flops_per_launch": flops},
            "valu_issue": valu_ref,
        },
        "roofline_batched": {
            "bound": "valu", "achieved": b_ach, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": b_ach / PEAK_FP32_TFLOPS, "traffic": b_traffic,
            "kernel": f"{mse_kernel}, one launch = the initial 1000 expressions x 1M",
            "kernel_ms": b_ms, "flops_per_launch": flops0,
            "hbm_algorithmic_bytes_per_launch": xy_bytes,
            "hbm_achieved_GBs": xy_bytes / (b_ms * 1e-3) / 1e9,
            "valu_issue": valu_mse,
        },
        "cpu_baseline": cpu,
        "cached_evolution": cached_gen,
        "host_planned_cached_evolution": host_cached,
        "batched_evolution": batched_gen,
        "plain_final_population": plainK,
        "batched_final_population": batchedK,
        "cached_over_plain": value / plainK["node_evals_per_s"],
        "cached_over_batched_generations": value / batched_gen["node_evals_per_s"],
        "plain_initial_population": plain0,
        "batched_initial_population": batched0,
        "batched_over_plain_initial": batched0["node_evals_per_s"] / plain0["node_evals_per_s"],
        "reference_cache_initial_population": ref_cache,
        "single_expression": c1,
        "multi_expression": c2,
        "mutant_cache": c3,
        "constant_fit": fit,
        "fp64_error": acc,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
