/*
 * sr_hip.h -- C ABI of the MI355X symbolic-regression evaluator (libsrhip.so).
 *
 * Plain C types only.  Every call returns an sr_status (0 = SR_OK); on error
 * sr_last_error(ctx) holds a message.  No call exits the process.
 *
 * Reference interfaces each entry point replaces (phi-jkim/symbolic_regression_gpu):
 *   sr_ctx_create / sr_ctx_destroy
 *       create_gpu_context / destroy_gpu_context      src/cleaned/cl_gpu_eval.cu:569-575
 *       create_gpu_simple_context / destroy_...        src/cleaned/cl_gpu_simple.cu:315-321
 *       create_gpu_nvrtc_context / destroy_...         src/cleaned/cl_gpu_nvrtc.cu:449-450
 *   sr_set_data / sr_set_data_device
 *       the upload_X flag + flatten_X_col_major_full   src/cleaned/cl_gpu_eval.cu:249-257,474-489
 *   sr_eval_mse (mode PLAIN / BATCHED / CACHED / JIT)
 *       evaluate_gpu_simple_wrapper                    src/cleaned/cl_gpu_simple.cu:324-336
 *       evaluate_gpu_mse_wrapper (subtree cache)       src/cleaned/cl_gpu_eval.cu:578-592
 *       evaluate_gpu_nvrtc_wrapper                     src/cleaned/cl_gpu_nvrtc.cu:452-548
 *       evaluate_cpu_mse (semantics)                   src/cleaned/cl_cpu_eval.cpp:94-168
 *   sr_eval_predictions
 *       MultiEvalFunc (predictions per expression)     src/utils/utils.h:141
 *   sr_cache_clear
 *       the clear_cache flag of evaluate_gpu_mse       src/cleaned/cl_gpu_eval.cu:344-347
 *   sr_detect
 *       detect_common_subtrees + mark_cached_subtrees_from_state
 *                                                      src/utils/detect.cpp:120-387,
 *                                                      src/cleaned/cl_gpu_eval.cu:294-329
 */
#ifndef SR_HIP_H
#define SR_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sr_ctx sr_ctx;

typedef enum {
    SR_OK = 0,
    SR_ERR_INVALID_ARG = -1,
    SR_ERR_HIP = -2,
    SR_ERR_BAD_EXPR = -3,  /* malformed token stream (underflow, bad token, bad var index) */
    SR_ERR_NO_DATA = -4,   /* no X|y uploaded */
    SR_ERR_CAPACITY = -5,  /* a fixed-capacity buffer would overflow */
    SR_ERR_JIT = -6,       /* hipRTC compile / load failure */
    SR_ERR_UNSUPPORTED = -7,
    SR_ERR_NO_DEVICE = -8,
    SR_ERR_IO = -9         /* file missing or unreadable / unwritable */
} sr_status;

typedef enum {
    SR_EVAL_PLAIN = 0,   /* one interpreter launch per expression */
    SR_EVAL_BATCHED = 1, /* one persistent length-ordered launch for the population */
    SR_EVAL_CACHED = 2,  /* batched + cross-population subtree cache */
    SR_EVAL_JIT = 3      /* hipRTC straight-line kernels */
} sr_eval_mode;

/* Packed prefix-token population (the layout of cl_gpu_simple.cu:213-227).
 * Expression e is tokens[offsets[e] .. offsets[e] + lengths[e]).
 * values: the doubles the reference InputInfo carries (hash fidelity); when
 * NULL, values_f32 is used. */
typedef struct {
    int32_t num_exprs;
    const int32_t *tokens;
    const double *values;
    const float *values_f32;
    const int32_t *offsets;
    const int32_t *lengths;
} sr_population;

/* Mirrors RunStats (utils.h:103-114) + EvalMetrics (utils.h:67-73). */
typedef struct {
    double total_eval_time_ms;
    double detect_time_ms;
    double transfer_time_ms;
    double kernel_time_ms;
    double jit_compile_time_ms;
    int32_t num_subtrees;
    double avg_subtree_size;
    double coverage;
    int32_t num_new_subtrees;
    int32_t num_cached;
    int32_t num_kernel_launches;
    int64_t logical_node_evals;  /* sum_e len_e * N */
    int64_t executed_node_evals; /* token evaluations actually run */
    double interp_kernel_ms;     /* interpreter kernels alone (HIP events on the context stream) */
    int32_t interp_launches;     /* interpreter kernel launches timed in interp_kernel_ms */
    int64_t executed_op_evals;   /* operator evaluations run (operator tokens * N; 0 = not counted) */
    int64_t column_reads;        /* cached-column operands read (datapoints) */
    int64_t column_writes;       /* cached-column values written (datapoints) */
    int32_t graph_captures;      /* hipGraphs captured by this call (sr_ga_run) */
} sr_stats;

/* Launch shape of the batched interpreter. */
typedef struct {
    int32_t dpt;           /* datapoints per lane: 1..4 (3 with <= 4 variables only) */
    int32_t chunk;         /* datapoints per work item (rounded up to a multiple of 768) */
    int32_t n_queues;      /* 1, or 8 = one queue per XCD */
    int32_t blocks_per_cu; /* persistent workgroups per CU (0 = occupancy limit) */
    int32_t cache_dpt;     /* datapoints per lane of cache-skeleton launches (0 = 3) */
    int32_t kernel;        /* whole expressions: 0 = lane interpreter (dpt datapoints per lane),
                              1 = vector interpreter (384 datapoints per instruction; default) */
    int32_t cache_kernel;  /* the same choice for cache skeletons (default 1) */
} sr_tuning;

int sr_ctx_create(int device, sr_ctx **out);
int sr_ctx_destroy(sr_ctx *ctx);
const char *sr_last_error(const sr_ctx *ctx);
/* HIP versions this library was built against (HIP_VERSION of its headers)
 * and runs on (hipRuntimeGetVersion / hipDriverGetVersion of the runtime the
 * process loaded; one libamdhip64.so.<major> serves the whole process, torch's
 * included).  SR_ERR_UNSUPPORTED when the runtime's major version differs
 * from the build's.  No counterpart in the reference. */
int sr_runtime_info(int32_t *built, int32_t *runtime, int32_t *driver);
/* Run on a caller-owned hipStream_t; NULL is the device's null (default)
 * stream, which is what torch's default stream reports as its handle. */
int sr_set_stream(sr_ctx *ctx, void *hip_stream);
/* Go back to the context's own non-blocking stream (the state after create). */
int sr_use_own_stream(sr_ctx *ctx);

/* Upload X|y, column-major [nv+1][N] fp32, y = column nv (host pointer). */
int sr_set_data(sr_ctx *ctx, const float *X_colmajor, int32_t nv, int64_t N);
/* Same layout, already in device memory; borrowed, caller keeps it alive. */
int sr_set_data_device(sr_ctx *ctx, const float *d_X_colmajor, int32_t nv, int64_t N);

/* Fitness: MSE over finite predictions (NaN when none), per expression.
 * mse_out: host [num_exprs]; valid_out (optional): host [num_exprs] finite
 * prediction counts; stats optional. */
int sr_eval_mse(sr_ctx *ctx, const sr_population *pop, int mode, float *mse_out,
                int64_t *valid_out, sr_stats *stats);
/* Allocation-free fitness.  No counterpart in the reference; for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  SR_EVAL_BATCHED on a
 * packed population held in DEVICE memory: d_tokens / d_values (fp32) of
 * num_tokens entries, d_offsets / d_lengths [num_exprs]; outputs d_mse and
 * d_valid (optional) [num_exprs] in device memory.  Only enqueues work on the
 * context stream (no synchronisation, no hipMalloc / hipFree, no memset
 * nodes): programs are compiled on the device, and every intermediate buffer
 * lives in the library's static device scratch of 64 MiB per device, which
 * bounds a call: programs take 8 B and compile scratch 28 B per token, MSE
 * partials 16 B per (expression, datapoint chunk).  A call whose layout does
 * not fit returns SR_ERR_CAPACITY.  The sum of the lengths must not exceed
 * num_tokens.  MSEs and counts equal sr_eval_mse (SR_EVAL_BATCHED) bit for
 * bit; an expression sr_eval_mse would reject (malformed, outside the token
 * arrays) gets MSE NaN and count 0 instead of an error.  Every call on a
 * device shares that device's scratch: calls that overlap on one device race. */
int sr_eval_mse_noalloc(sr_ctx *ctx, int32_t num_exprs, int64_t num_tokens, const int32_t *d_tokens,
                        const float *d_values, const int32_t *d_offsets, const int32_t *d_lengths, float *d_mse,
                        int64_t *d_valid);
/* Predictions, host [num_exprs][N]. */
int sr_eval_predictions(sr_ctx *ctx, const sr_population *pop, float *pred_out);

/* ---- subtree cache (mode SR_EVAL_CACHED) --------------------------------
 * capacity: value columns kept on the device (reference MAX_CACHED_SUBTREES
 * 2048, cl_gpu_common.h:13; each column is N floats); min_size / min_freq:
 * detection thresholds (reference SR system: 3 / 10, cl_gpu_eval.cu:371-373).
 * No eviction: once full, further subtrees are not cached (the reference's
 * policy).  The cache persists across calls until sr_cache_clear or new data. */
int sr_cache_config(sr_ctx *ctx, int32_t capacity, int32_t min_size, int32_t min_freq);
int sr_cache_clear(sr_ctx *ctx);
/* Cache policy for SR_EVAL_CACHED.
 *  SR_CACHE_REFERENCE (default): the reference's policy above.
 *  SR_CACHE_GENERATIONAL, SR_CACHE_GENERATIONAL_REUSED,
 *  SR_CACHE_GENERATIONAL_SHARED: no counterpart in
 *    the reference (its cache keeps what it first detected and never evicts,
 *    cl_gpu_eval.cu:294-329).  The cache follows the population: a subtree
 *    (size >= min_size) with a column is read instead of evaluated; a column
 *    is written, inside the launch that evaluates the subtree anyway, for
 *    every new subtree (GENERATIONAL), for a subtree met twice -- twice in
 *    one population or in two consecutive ones (GENERATIONAL_REUSED), or
 *    for a subtree met twice in one population and for a subtree nested in
 *    one of those that the previous population also held
 *    (GENERATIONAL_SHARED: a subtree met once is rarely met again, so its
 *    column is seldom read);
 *    columns of subtrees absent from the population being evaluated are
 *    evicted; an expression evaluated before keeps its MSE.  capacity 0
 *    sizes the columns from free device memory (half of it, at most 131072
 *    columns of N floats).  MSEs and counts equal SR_EVAL_BATCHED's bit for
 *    bit (hits are verified against the stored token body).
 * Any call empties the cache. */
enum {
    SR_CACHE_REFERENCE = 0,
    SR_CACHE_GENERATIONAL = 1,
    SR_CACHE_GENERATIONAL_REUSED = 2,
    SR_CACHE_GENERATIONAL_SHARED = 3,
    /* The generational cache planned on the device, for the context's GA
     * population (sr_ga_run, sr_ga_evaluate): every subtree of >= min_size
     * tokens with a column is read, every subtree executed at least twice in
     * the population is evaluated once and its column kept, columns of
     * subtrees the population no longer holds are evicted; no host work, so
     * sr_ga_run with use_graph replays one captured graph per generation.
     * sr_eval_mse on a host population does not take this policy. */
    SR_CACHE_DEVICE = 4
};
int sr_cache_policy(sr_ctx *ctx, int32_t policy, int32_t capacity, int32_t min_size);

/* Host planner of the generational cache, without a device (inspection and
 * tests).  plan: validated streams (as sr_eval_mse) -> launches of skeleton
 * programs: OP_REF (-2, value = slot) reads a column written by an earlier
 * launch, OP_STORE (-3, value = slot) evaluates its operand subtree and also
 * writes that column.  plan_copy (any pointer may be NULL): the programs
 * packed (num_tokens tokens/values, num_programs offsets/lengths), launch k =
 * programs [launch_begin[k], launch_begin[k+1]) (num_launches + 1 entries),
 * root_slot per program (slot holding its whole value, or -1), expr_src per
 * expression (program giving its MSE, or -2 - slot of a remembered MSE).
 * commit: MSE and finite count per program, after running the launches. */
typedef struct sr_memo_planner sr_memo_planner;
int sr_memo_planner_create(int32_t capacity, int32_t min_size, int32_t policy, sr_memo_planner **out);
int sr_memo_planner_destroy(sr_memo_planner *p);
int sr_memo_plan(sr_memo_planner *p, const sr_population *pop, int32_t *num_launches,
                 int32_t *num_programs, int64_t *num_tokens);
int sr_memo_plan_copy(const sr_memo_planner *p, int32_t *tokens, double *values, int32_t *offsets,
                      int32_t *lengths, int32_t *launch_begin, int32_t *root_slot, int32_t *expr_src);
int sr_memo_commit(sr_memo_planner *p, const float *mse, const int64_t *cnt);
int sr_memo_remembered(const sr_memo_planner *p, int32_t slot, float *mse, int64_t *cnt);
/* Flat hashes of the cached subtrees, by slot. */
int sr_cache_state(sr_ctx *ctx, uint64_t *keys_out, int32_t max_keys, int32_t *num_cached);
/* Cache marking of the last CACHED evaluation, packed like its population:
 * slot at a cached subtree's root, -2 inside it, -1 elsewhere. */
int sr_cache_last_hints(sr_ctx *ctx, int32_t *hints_out, int64_t max_len, int64_t *len_out);
/* Stateless common-subtree detection: hints packed like the population
 * (subtree id at an occurrence root, -2 inside, -1 elsewhere); per detected
 * subtree (size descending) its flat hash, size and occurrence count. */
int sr_detect(sr_ctx *ctx, const sr_population *pop, int32_t min_size, int32_t min_freq,
              int32_t *hints_out, uint64_t *flat_out, int32_t *size_out, int32_t *count_out,
              int32_t max_subs, int32_t *num_subs);

/* ---- constant fitting ------------------------------------------------------
 * No counterpart in the reference (it has no gradient code); these serve the
 * differentiable-fitting callers (the torch.autograd.Function shim).
 * const_slot: per token of pop, -1 (fixed) or the free-constant slot 0..7 of
 * its expression (CONST tokens only; tokens sharing a slot share a value).
 * init_vals (optional, host [E*8]): the population's values of each slot.
 * The eval / grad calls take DEVICE pointers and only enqueue work on the
 * context stream (sr_set_stream); sr_sync waits for it.
 *   cvals [E*8] constants, mse [E], grad [E*8] = dMSE/dconstant. */
int sr_param_prepare(sr_ctx *ctx, const sr_population *pop, const int32_t *const_slot,
                     float *init_vals);
int sr_param_info(sr_ctx *ctx, int32_t *num_exprs, int32_t *nslots_out, int64_t *tokens);
/* Forward MSE with the given constants. */
int sr_param_eval(sr_ctx *ctx, const float *d_cvals, float *d_mse);
/* Reverse-mode adjoint: MSE and dMSE/dconstants in one launch. */
int sr_param_grad(sr_ctx *ctx, const float *d_cvals, float *d_mse, float *d_grad);
/* Central differences (step h) on the GPU: 2 forward passes per used slot. */
int sr_param_fd_grad(sr_ctx *ctx, const float *d_cvals, float h, float *d_grad);
int sr_sync(sr_ctx *ctx);
/* The gradient entry point of the drop-in boundary on host arrays: dMSE/d
 * constant at the population's own values for every CONST token whose
 * const_mask byte is nonzero (each its own parameter, at most 8 per
 * expression), by the reverse-mode adjoint above; synchronous.  grad_out is
 * packed like pop->tokens (0 where the mask is 0); mse_out optional [E]. */
int sr_grad_consts(sr_ctx *ctx, const sr_population *pop, const uint8_t *const_mask, float *grad_out,
                   float *mse_out);

/* Compile one expression to the batched interpreter's bytecode (host only, no
 * GPU needed).  out_code/out_arg: [len]; returns the instruction count, or a
 * negative sr_status.  *stack_need = per-lane stack slots the program uses.
 * A test and inspection hook; no counterpart in the reference. */
int sr_compile_expr(const int32_t *tokens, const float *values, int32_t len, int32_t nv,
                    int32_t *out_code, int32_t *out_arg, int32_t *stack_need);
/* Compile one cache skeleton (OP_REF = -2 and OP_STORE = -3 tokens accepted,
 * OP_STORE's value = destination column) with algorithm 0 = the serial
 * compiler of sr_compile_expr, 1 = the wave-parallel compiler the
 * device-planned cache runs on the GPU, here executed on the host lane by
 * lane.  Both give the same bytecode, the first four OP_REF operands carrying
 * 1 + their order in code bits 16..18.  Returns the instruction count or a
 * negative sr_status.  A test and inspection hook; no counterpart in the
 * reference. */
int sr_compile_skeleton(const int32_t *tokens, const float *values, int32_t len, int32_t nv, int32_t algorithm,
                        int32_t *out_code, int32_t *out_arg, int32_t *stack_need);
/* Same, as the parametric program the adjoint pass runs: const_slot [len] is
 * -1 or the free-constant slot (0..7) of each token.  Code bit 15 marks
 * instructions whose value depends on a free constant, bits 20..30 hold 1 +
 * the tape row of a value the reverse pass reads; a SRC_POP's arg is the tape
 * row of the popped value.  out_aux (optional) [len]: for an active
 * instruction the tape row of the previous top when it is an operand (else
 * -1), for an inactive one the first instruction of its subtree.
 * *tape_rows (optional): tape rows the program needs. */
int sr_compile_param_expr(const int32_t *tokens, const float *values, const int32_t *const_slot,
                          int32_t len, int32_t nv, int32_t *out_code, int32_t *out_arg, int32_t *out_aux,
                          int32_t *stack_need, int32_t *tape_rows);

/* hipRTC backend (mode SR_EVAL_JIT; replaces the NVRTC path of
 * cl_gpu_nvrtc.cu:152-548).  Host-only helpers, no GPU needed: the generated
 * HIP source for a population (nv variables, dpt datapoints per lane) and a
 * compile of it for gfx950 (log, compile ms, code-object bytes).  sr_eval_mse
 * in mode SR_EVAL_JIT compiles once per distinct population (stats report
 * jit_compile_time_ms next to kernel_time_ms). */
int sr_jit_source(const sr_population *pop, int32_t nv, int32_t dpt, char *out, int64_t cap,
                  int64_t *len);
int sr_jit_compile(const sr_population *pop, int32_t nv, int32_t dpt, char *log_out, int64_t log_cap,
                   double *compile_ms, int64_t *code_bytes);
/* The compiled code object itself (out may be NULL to query *len). */
int sr_jit_code(const sr_population *pop, int32_t nv, int32_t dpt, void *out, int64_t cap, int64_t *len);

/* Autotuner (no counterpart in the reference): times every valid launch
 * shape of the batched interpreter (lane interpreter with 1..4 datapoints per
 * lane or vector interpreter, chunk = datapoints per work item, XCD queues,
 * workgroups per CU) on a length-stratified sample of at most max_sample
 * expressions of pop (0: 256), sets the fastest as the context tuning and
 * remembers it per (architecture, CU count, variable-register class, N).
 * best_ms: kernel ms of the winner on the sample (0 when reused). */
int sr_autotune(sr_ctx *ctx, const sr_population *pop, int32_t max_sample, sr_tuning *best, double *best_ms);
/* Skeleton kernel of the device-planned subtree cache (cache_kernel,
 * cache_dpt): re-runs both interpreter launches of the last cached
 * generation (SR_CACHE_DEVICE; identical columns and partial sums) with every
 * candidate, keeps the fastest in the context tuning and the remembered
 * entry, and drops captured generation graphs.  best_ms: the winner's ms for
 * both launches.  SR_ERR_INVALID_ARG before a cached generation ran on the
 * current data. */
int sr_autotune_cache(sr_ctx *ctx, sr_tuning *best, double *best_ms);
/* The remembered tunings as text, one entry per line
 * ("arch cus xc N dpt chunk n_queues blocks_per_cu cache_dpt kernel
 * cache_kernel"); sr_tune_load adds a file's entries (all or none: a bad line
 * fails the call with SR_ERR_INVALID_ARG), so a later sr_autotune of the same
 * shape reuses them without timing.  entries may be NULL. */
int sr_tune_save(const char *path);
int sr_tune_load(const char *path, int32_t *entries);
/* Same for the constant-fitting kernels of the prepared population: picks
 * datapoints per lane for the forward (finite differences) and adjoint
 * kernels at the given constants (device pointer [E*8]). */
int sr_param_autotune(sr_ctx *ctx, const float *d_cvals, int32_t *fwd_dpt, int32_t *adj_dpt);

int sr_get_tuning(sr_ctx *ctx, sr_tuning *out);
int sr_set_tuning(sr_ctx *ctx, const sr_tuning *t);

/* ---- genetic algorithm on a device-resident population ------------------
 * Replaces the host GA of the reference: generate_random_expression /
 * mutate_expression / crossover_expressions (src/utils/utils.cpp:1156-1272),
 * the selection loop of src/cleaned/cl_sr_system.cpp:285-330, and the evogp
 * device kernels generate / mutate / crossover (src/utils/gpu_kernel.h:15-58).
 * Randomness: Philox4x32-10 streams keyed by (seed, individual, generation),
 * so offspring depend only on (seed, generation, population, fitness). */
typedef struct {
    int32_t num_vars;
    int32_t max_tokens;  /* token budget of generate / mutate / crossover */
    int32_t max_depth;
    float const_min, const_max;
    float prob_leaf, prob_const_leaf, prob_unary;
    int32_t n_unary, n_binary;
    int32_t unary_ops[32];
    int32_t binary_ops[32];
} sr_ga_params;

typedef struct {
    int32_t tournament; /* reference: 3 */
    float p_mutation;   /* reference: 0.2 (rest crossover) */
    double elite_frac;  /* reference: 0.05, at least one elite */
} sr_select_params;

/* Reference defaults (utils.h:178-187, utils.cpp:938-947). */
int sr_ga_params_default(sr_ga_params *p);
/* New random population of pop_size on the device (stream domain "init").
 * Individuals are stored with a fixed stride (max_tokens + max_depth + 2). */
int sr_ga_init(sr_ctx *ctx, const sr_ga_params *p, int32_t pop_size, uint64_t seed);
/* Copy the population out: tok/val [pop*cap] (stride cap), len [pop].
 * Any pointer may be NULL; pop/cap are always written when non-NULL. */
int sr_ga_get_population(sr_ctx *ctx, int32_t *tok, float *val, int32_t *len, int32_t *pop,
                         int32_t *cap);
/* Replace the population (stride cap must equal the context's). */
int sr_ga_set_population(sr_ctx *ctx, const int32_t *tok, const float *val, const int32_t *len,
                         int32_t pop, int32_t cap);
/* Fitness of the current population (any sr_eval_mode); mse_out host [pop]. */
int sr_ga_evaluate(sr_ctx *ctx, int mode, float *mse_out, sr_stats *stats);
/* Next generation from a fitness vector (host [pop]; NULL = the last
 * sr_ga_evaluate result), generation index gen. */
int sr_ga_next(sr_ctx *ctx, const sr_select_params *sp, const float *fitness, uint64_t seed,
               uint32_t gen);
/* gens generations of evaluate -> fitness -> select -> vary, generation
 * indices gen0 .. gen0+gens-1.  mode BATCHED with use_graph=1 replays one
 * captured hipGraph per generation.  mse_hist (optional): host [gens*pop]
 * fitness of each evaluated generation.  ms_out (optional): wall ms. */
int sr_ga_run(sr_ctx *ctx, const sr_select_params *sp, int mode, int32_t gens, uint32_t gen0,
              uint64_t seed, int use_graph, float *mse_hist, double *ms_out, sr_stats *stats);
/* One select + vary step on host arrays, the evolve step of the drop-in
 * boundary (cl_sr_system.cpp:285-330 applied to a caller's population):
 * packed pop_in (validated) and fitness [E] -> the next generation, stride
 * cap_out = max_tokens + max_depth + 2: tok_out / val_out [E*cap_out],
 * len_out [E].  Equals sr_ga_set_population + sr_ga_next + sr_ga_get_population
 * and replaces the context's GA population. */
int sr_evolve_step(sr_ctx *ctx, const sr_population *pop_in, const float *fitness, const sr_ga_params *p,
                   const sr_select_params *sp, uint64_t seed, uint32_t gen, int32_t cap_out, int32_t *tok_out,
                   float *val_out, int32_t *len_out);

/* Batch subtree replacement on the device (the evogp mutate / crossover ABI,
 * src/utils/gpu_kernel.h:31-58, on prefix tokens): child k is left tree
 * lsel[k] with the subtree at node lidx[k] replaced by the subtree of right
 * tree rsel[k] at node ridx[k].  Trees are stored with stride cap in host
 * arrays; a child that would exceed cap, or an index outside its tree, is a
 * copy of the left tree.  Outputs: otok/oval [n*cap], olen [n]. */
int sr_tree_splice(sr_ctx *ctx, int32_t n, int32_t cap, const int32_t *ltok, const float *lval,
                   const int32_t *llen, int32_t n_left, const int32_t *lsel, const int32_t *lidx,
                   const int32_t *rtok, const float *rval, const int32_t *rlen, int32_t n_right,
                   const int32_t *rsel, const int32_t *ridx, int32_t *otok, float *oval,
                   int32_t *olen);

/* ---- evogp population layout (defs.h:30-107, gpu_kernel.h:15-58) ---------
 * Per individual, max_len slots of value (float: constant, variable index or
 * opcode), node type (int16: VAR 0, CONST 1, UFUNC 2, BFUNC 3; bit 7, the
 * output flag, is ignored) and subtree size (int16; size[0] = tree length),
 * in prefix order.  Device pointers; the context's stream, synchronous.
 * evogp -> prefix restates gpu_simple_jinha_with_evolve.cu:846-866 (a
 * function node's value is its opcode); an individual whose size[0] is
 * outside 1..min(max_len, cap) gets length 0.  prefix -> evogp computes the
 * node types from token arity and the subtree sizes; a stream longer than
 * max_len or malformed gets size[0] = 0.  The GA forms read / write the
 * context's device population (sr_ga_init sets its stride). */
int sr_evogp_to_prefix(sr_ctx *ctx, int32_t pop, int32_t max_len, const float *d_value, const int16_t *d_type,
                       const int16_t *d_size, int32_t cap, int32_t *d_tok, float *d_val, int32_t *d_len);
int sr_prefix_to_evogp(sr_ctx *ctx, int32_t pop, int32_t cap, const int32_t *d_tok, const float *d_val,
                       const int32_t *d_len, int32_t max_len, float *d_value, int16_t *d_type, int16_t *d_size);
int sr_ga_import_evogp(sr_ctx *ctx, int32_t pop, int32_t max_len, const float *d_value, const int16_t *d_type,
                       const int16_t *d_size);
int sr_ga_export_evogp(sr_ctx *ctx, int32_t max_len, float *d_value, int16_t *d_type, int16_t *d_size);

/* ---- on-disk formats (host only, no device needed) -----------------------
 * Digest (gen_k.txt) as parse_input_info reads it (utils.cpp:122-233): the
 * expression count, then per expression "num_vars num_dps num_tokens", the
 * tokens, the values and the data-file path on its own line; shared_data = 1
 * when there are several expressions and all name the same file, variable and
 * datapoint count (utils.cpp:202-223).  Numbers parse as the reference's
 * stream reads do; a truncated or malformed digest is SR_ERR_BAD_EXPR. */
typedef struct sr_digest sr_digest;
int sr_digest_read(const char *path, sr_digest **out);
int sr_digest_free(sr_digest *d);
int sr_digest_shape(const sr_digest *d, int32_t *num_exprs, int64_t *num_tokens, int32_t *shared_data);
/* Packed like sr_population (offsets into tokens / values); any pointer may
 * be NULL. */
int sr_digest_population(const sr_digest *d, int32_t *tokens, double *values, int32_t *offsets,
                         int32_t *lengths, int32_t *num_vars, int32_t *num_dps);
int sr_digest_data_path(const sr_digest *d, int32_t expr, char *out, int64_t cap, int64_t *len);
/* The same format, values printed with 17 significant digits (exact). */
int sr_digest_write(const char *path, const sr_population *pop, const int32_t *num_vars,
                    const int32_t *num_dps, const char *const *data_paths);
/* Data file as load_data_file reads it (utils.cpp:235-300): num_dps rows of
 * num_vars + 1 numbers (x..., y) parsed with strtod; out is column-major
 * double [num_vars+1][num_dps] (the reference's vars[i][j]); parsed = numbers
 * actually converted (a short file leaves zeros, as in the reference). */
int sr_data_file_read(const char *path, int32_t num_vars, int64_t num_dps, double *out, int64_t *parsed);
int sr_data_file_write(const char *path, int32_t num_vars, int64_t num_dps, const double *cols);

#ifdef __cplusplus
}
#endif
#endif
