"""Replay of the reference's on-disk evolution traces and its per-generation
CSV, on this evaluator.

* ``replay`` restates ``evaluate_evolution_benchmark`` (utils.cpp:1275-1393):
  generations ``gen_<start>.txt`` .. ``gen_<end>.txt`` of a directory are
  parsed (``read_digest``), the shared data file is loaded once (from the
  first generation, when every expression names it), each generation is
  evaluated with the chosen mode (data loading is outside the timed region,
  as in the reference) and one summary line per generation is printed in the
  reference's wording.  Digests whose expressions name different data files
  are evaluated file by file.
* ``generation_row`` / ``CSV_HEADER`` restate the SR system's CSV
  (cl_sr_system.cpp:133-145,229-283): best and median MSE over the finite
  MSEs, token totals, timing and cache statistics per generation.
* ``evolve`` runs the device GA (evaluate -> fitness -> select -> vary) for a
  number of generations and writes that CSV; ``write_trace`` writes the
  populations it visits as a replayable trace (``gen_<k>.txt`` +
  ``shared_data.txt``, the layout generate_evolution_test20.py:104-170
  writes).

Command line:  python -m symbolic_regression_gpu_amd.replay DIR --start 0
--end 19 [--mode cached|batched|plain] [--csv out.csv] [--data-root ROOT]
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

from .api import (CACHE_GENERATIONAL_SHARED, EVAL_BATCHED, EVAL_CACHED, EVAL_PLAIN, Context,
                  Population, SrError, read_data_file, read_digest, write_data_file, write_digest)

CSV_HEADER = ("Gen,BestMSE,MedianMSE,TotalTokens,CoveredTokens,TotalTimeMs,DetectTimeMs,"
              "H2D_D2H_TimeMs,JITTimeMs,KernelTimeMs,NumSubtrees,AvgSubtreeSize,Coverage")
MODES = {"plain": EVAL_PLAIN, "batched": EVAL_BATCHED, "cached": EVAL_CACHED}


def generation_row(gen: int, mses, lengths, stats: dict) -> dict:
    """One CSV row (cl_sr_system.cpp:229-283): best = smallest finite MSE
    (1e9 when none), median = the middle of the sorted finite MSEs (0 when
    none), covered tokens = total tokens x coverage."""
    m = np.asarray(mses, np.float32)
    fin = np.sort(m[np.isfinite(m)])
    total = float(np.sum(lengths))
    return {
        "Gen": gen,
        "BestMSE": float(fin[0]) if fin.size else 1e9,
        "MedianMSE": float(fin[fin.size // 2]) if fin.size else 0.0,
        "TotalTokens": total,
        "CoveredTokens": total * stats.get("coverage", 0.0),
        "TotalTimeMs": stats.get("total_eval_time_ms", 0.0),
        "DetectTimeMs": stats.get("detect_time_ms", 0.0),
        "H2D_D2H_TimeMs": stats.get("transfer_time_ms", 0.0),
        "JITTimeMs": stats.get("jit_compile_time_ms", 0.0),
        "KernelTimeMs": stats.get("kernel_time_ms", 0.0),
        "NumSubtrees": stats.get("num_subtrees", 0),
        "AvgSubtreeSize": stats.get("avg_subtree_size", 0.0),
        "Coverage": stats.get("coverage", 0.0),
    }


def _csv_line(row: dict) -> str:
    return ",".join(str(row[k]) for k in CSV_HEADER.split(","))


def _subset(pop: Population, ids) -> Population:
    exprs = [pop.expr(int(e)) for e in ids]
    return Population.from_exprs(exprs)


def replay(data_dir: str, start: int, end: int, mode: int = EVAL_CACHED, ctx: Context | None = None,
           csv_path: str | None = None, data_root: str | None = None, log=print) -> list:
    """Evaluate generations start..end of a trace directory; returns one dict
    per evaluated generation (its MSEs, valid counts, wall ms, summary, CSV
    row).  Relative data paths resolve against data_root (default: the
    current directory, as the reference's binary does)."""
    ctx = ctx or Context(0)
    root = data_root if data_root is not None else os.getcwd()

    def resolve(p):
        return p if os.path.isabs(p) else os.path.join(root, p)

    log(f"Running Evolution Benchmark: Gen {start} to {end}")
    loaded = None
    total_ms = 0.0
    out = []
    csv = open(csv_path, "w") if csv_path else None
    if csv:
        csv.write(CSV_HEADER + "\n")
    try:
        for gen in range(start, end + 1):
            path = os.path.join(data_dir, f"gen_{gen}.txt")
            try:
                d = read_digest(path)
            except SrError:
                log(f"Skipping generation {gen} (failed to load)")
                continue
            pop = d.population
            E = pop.num_exprs
            groups = {}
            for e in range(E):
                groups.setdefault((d.data_paths[e], int(d.num_vars[e]), int(d.num_dps[e])), []).append(e)
            mse = np.empty(E, np.float32)
            valid = np.empty(E, np.int64)
            stats = {}
            dt = 0.0
            for key, ids in groups.items():
                if key != loaded:
                    cols, _ = read_data_file(resolve(key[0]), key[1], key[2])
                    ctx.set_data(cols.astype(np.float32))
                    loaded = key
                sub = pop if len(groups) == 1 else _subset(pop, ids)
                t0 = time.perf_counter()
                m, v, st = ctx.eval_mse(sub, mode)
                dt += 1e3 * (time.perf_counter() - t0)
                mse[ids] = m
                valid[ids] = v
                for k, x in st.items():
                    stats[k] = stats.get(k, 0) + x if isinstance(x, (int, float)) else x
            if len(groups) > 1:  # cache statistics describe the last group only
                stats["coverage"] = st.get("coverage", 0.0)
            total_ms += dt
            fin = mse[np.isfinite(mse)]
            min_mse = float(fin.min()) if fin.size else float("nan")
            avg_mse = float(np.mean(mse.astype(np.float64)))
            summary = f"Gen {gen}: {dt:g} ms | Min MSE: {min_mse:g} | Avg MSE: {avg_mse:g}"
            log(summary)
            row = generation_row(gen, mse, pop.lengths, stats)
            if csv:
                csv.write(_csv_line(row) + "\n")
            out.append({"gen": gen, "mse": mse, "valid": valid, "ms": dt, "summary": summary, "row": row})
    finally:
        if csv:
            csv.close()
    log(f"### Total Evolution Time: {total_ms:g} ms ###")
    return out


def evolve(ctx: Context, params, sel, gens: int, seed: int, pop_size: int, mode: int = EVAL_CACHED,
           csv_path: str | None = None, trace_dir: str | None = None, data_cols=None, log=print) -> list:
    """Device GA for `gens` generations from a fresh population (the SR
    system's loop, cl_sr_system.cpp:168-333), one CSV row per generation;
    with trace_dir, each evaluated population is written as gen_<k>.txt next
    to shared_data.txt (data_cols = the X|y columns set on ctx)."""
    ctx.ga_init(params, pop_size, seed)
    rows = []
    csv = open(csv_path, "w") if csv_path else None
    if csv:
        csv.write(CSV_HEADER + "\n")
    if trace_dir is not None:
        os.makedirs(trace_dir, exist_ok=True)
        if data_cols is None:
            raise ValueError("write_trace needs the data columns")
        write_data_file(os.path.join(trace_dir, "shared_data.txt"), np.asarray(data_cols, np.float64))
    try:
        for g in range(gens):
            tok, val, ln, cap = ctx.ga_population()
            m, st = ctx.ga_evaluate(mode)
            if trace_dir is not None:
                pop = Population.from_strided(tok, val, ln, cap)
                nv = int(np.asarray(data_cols).shape[0] - 1)
                N = int(np.asarray(data_cols).shape[1])
                write_digest(os.path.join(trace_dir, f"gen_{g}.txt"), pop, nv, N,
                             os.path.join(trace_dir, "shared_data.txt"))
            row = generation_row(g, m, ln, st)
            rows.append(row)
            if csv:
                csv.write(_csv_line(row) + "\n")
            log(f"Gen {g}: Best MSE: {row['BestMSE']:g} Time: {st.get('total_eval_time_ms', 0.0):g}ms")
            ctx.ga_next(sel, m, seed, g)
    finally:
        if csv:
            csv.close()
    return rows


def write_trace(ctx: Context, trace_dir: str, params, sel, gens: int, seed: int, pop_size: int,
                data_cols, mode: int = EVAL_BATCHED) -> list:
    """A replayable trace of `gens` device-GA generations (see evolve)."""
    return evolve(ctx, params, sel, gens, seed, pop_size, mode, trace_dir=trace_dir, data_cols=data_cols,
                  log=lambda *a: None)


def main(argv=None):
    ap = argparse.ArgumentParser(description="replay gen_<k>.txt traces on the MI355X evaluator")
    ap.add_argument("data_dir")
    ap.add_argument("--start", type=int, default=0)
    ap.add_argument("--end", type=int, default=19)
    ap.add_argument("--mode", choices=sorted(MODES), default="cached")
    ap.add_argument("--csv", default=None)
    ap.add_argument("--data-root", default=None)
    args = ap.parse_args(argv)
    ctx = Context(0)
    if args.mode == "cached":
        ctx.cache_policy(CACHE_GENERATIONAL_SHARED, 0, 3)
    replay(args.data_dir, args.start, args.end, MODES[args.mode], ctx, args.csv, args.data_root)
    return 0


if __name__ == "__main__":
    sys.exit(main())
