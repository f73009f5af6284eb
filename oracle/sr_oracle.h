/*
 * sr_oracle.h -- CPU restatement of the reference symbolic-regression hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the oracle of the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  Nothing in symbolic_regression_gpu_amd/ links or calls it.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Pinning status per function (what checks it independently of this code):
 *   sro_op32 / sro_eval_point32 / sro_predict32 / sro_mse32
 *       pinned: golden vectors from the reference evaluator compiled in place
 *       (oracle/ref_harness.cpp, cl_cpu_eval.cpp:16-168) + closed forms.
 *   sro_op64 / sro_eval_point64 / sro_predict64
 *       pinned: reference fp64 op table compiled in place
 *       (cpu_simple_single.cpp:6-79) for the ops both share; closed forms.
 *   sro_hash_flat / sro_subtree_merkle
 *       pinned: golden u64 values from detect.cpp:50-118 compiled in place.
 *   sro_detect
 *       pinned: detect_common_subtrees (detect.cpp:120-387) compiled in place,
 *       compared as hash sets / occurrence sets / hint maps modulo ids.
 *   sro_mark_cached / sro_cache_update / sro_skeleton / sro_predict32_cached
 *       the reference implementation (cl_gpu_eval.cu:165-186,294-329,384-397,
 *       433-459) is CUDA source that cannot be compiled here.  Pinned by
 *       tests/test_oracle_pins.py against brute-force models written there:
 *       hint maps equal the unique outermost-first marking of cached bodies
 *       (every subtree compared with every cached body); new slots follow the
 *       detected-subtree order up to capacity, skipping cached bodies (forced
 *       flat-hash collisions included: slots compare bodies as well as flat
 *       hashes, the reference's flat-hash-only keys collide); a skeleton
 *       expanded by its slots' bodies is the original stream, and evaluated
 *       over predict32 columns it equals predict32 of the expression bit for
 *       bit.
 *   sro_subtree_predict64
 *       pinned: eval_cpu_common_subtree_batch (cpu_subtree.cpp:135-340)
 *       compiled in place, FP64 predictions bit for bit; equal to
 *       sro_predict64 bit for bit.
 *   sro_subtree_state_*
 *       pinned by invariants: equal to sro_predict64 bit for bit over
 *       generations, new-entry counts equal a brute-force count.  The
 *       reference's eval_cpu_stateful_batch keys subtrees by the wrong span
 *       (see below), so it is not a golden source.
 *   sro_philox4x32_10
 *       pinned: Random123 published known-answer vectors.
 *   sro_ga_generate / sro_ga_mutate / sro_ga_crossover
 *       pinned: with the RNG hooks bound to std::mt19937 + libstdc++
 *       distributions, outputs equal the reference's
 *       generate_random_expression / mutate_expression / crossover_expressions
 *       (utils.cpp:986-1272) compiled in place, bit for bit.
 *   sro_fitness_order / sro_ga_next_generation(_hooked) (sort + elitism +
 *       tournament + variation choice, cl_sr_system.cpp:285-330; the reference
 *       code sits inside main() and cannot be called): pinned by
 *       tests/test_oracle_pins.py with scripted draws: order = a stable sort
 *       with NaN last; elites = the best max(1, floor(pop * frac)) copied;
 *       per child t draws in [0, pop), strict < keeps the earlier draw (a NaN
 *       drawn first is never displaced, as in the reference), mutation iff the
 *       next uniform draw < p_mutation, a second tournament only for
 *       crossover, and the child equals the pinned mutate / crossover replayed
 *       on the remaining draws with those parents.
 *   sro_op32 for SR_OP_EXP_CLAMPED (a new opcode 29, not in the reference
 *       evaluator) follows the clamped exp of evolution_tests.cpp:40; parity
 *       for it is closed-form only.
 *   sro_predict64_bound (FP64 value + first-order error bound of an fp32
 *       evaluation; no reference counterpart): pinned by the fp32 oracle
 *       lying inside it on every point of all-opcode populations, by planted
 *       ulp errors being caught, and by exact / branch-flip special cases.
 *   sro_mse_grad64_bound (FP64 forward-mode dMSE/dconstants + error bound of
 *       an fp32 adjoint; no reference counterpart, the reference has no
 *       gradient code): the gradient is pinned to torch FP64 autograd; the
 *       bound by a fp32 emulation of the adjoint pass staying inside it and a
 *       1% error planted in any opcode's partial landing far outside it.
 */
#ifndef SR_ORACLE_H
#define SR_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SRO_TOK_CONST 0
#define SRO_TOK_VAR (-1)
#define SRO_OP_REF (-2)
#define SRO_OP_EXP_CLAMPED 29
#define SRO_MAX_STACK 512

/* ---- op semantics + interpreter (cl_cpu_eval.cpp:16-90) ---- */
float sro_op32(int op, float a, float b);
double sro_op64(int op, double a, double b);
int sro_is_binary(int tok); /* interpreter arity rule: 1..9 binary */

float sro_eval_point32(const int *tok, const double *val, int len, const float *x, int nv);
double sro_eval_point64(const int *tok, const double *val, int len, const double *x, int nv);

/* Column-major X[v*N + dp], v < nv; y lives in column nv (not read here). */
void sro_predict32(const int *tok, const double *val, int len, const float *X, int nv,
                   int64_t N, float *out);
void sro_predict64(const int *tok, const double *val, int len, const float *X, int nv,
                   int64_t N, double *out);

/* Packed population: expression e = tok[off[e] .. off[e]+len[e]).
 * MSE per cl_cpu_eval.cpp:94-168: fp32 pred, fp32 diff*diff, fp64 sum over
 * finite preds, NaN when none.  threads<=0 -> 1.  Also returns valid count. */
void sro_mse32(const int *tok, const double *val, const int *off, const int *len, int E,
               const float *X, int nv, int64_t N, int threads, float *mse_out,
               int64_t *valid_out);
/* FP64 evaluator MSE (pred in fp64, diff^2 in fp64) -- error yardstick. */
void sro_mse64(const int *tok, const double *val, const int *off, const int *len, int E,
               const float *X, int nv, int64_t N, int threads, double *mse_out);

/* FP64 evaluation with a first-order running error bound (forward error
 * analysis of an fp32 evaluation): every fp32 operation adds its rounding
 * (0.5 ulp for correctly rounded ops, 2 ulp for libm transcendentals) and
 * propagates its operands' bounds through |df/da|, |df/db|.  bound_out is
 * +inf where a protected branch (DIV/LOOSE_* guards, clamped exp, log/sqrt
 * domain, asin/acos domain) could flip within the bound, or where an
 * intermediate leaves the fp32 range ambiguously.  Values beyond the fp32
 * range become +-inf as in fp32.  Not pinned by the reference: it is the
 * yardstick for the parity tolerance; checked against fp32 evaluations. */
/* FP64 MSE and dMSE/dconstant over the points where the fp32 oracle is finite,
 * with constants of slot k (slot[i] = -1 or 0..7) set to cvals[k], and a bound
 * on the error of an fp32 adjoint evaluation of each gradient component
 * (operand rounding carried through the partials, fp32 summation of
 * lane_terms points per lane).  bound = inf where rounding can flip a branch. */
void sro_mse_grad64_bound(const int *tok, const double *val, const int *slot, int len, const float *X,
                          int nv, int64_t N, const float *cvals, int lane_terms, double *mse_out,
                          double *grad_out, double *bound_out, int64_t *count_out);
void sro_predict64_bound(const int *tok, const double *val, int len, const float *X, int nv,
                         int64_t N, double *out, double *bound_out);

/* Maximum stack depth the right-to-left interpreter reaches; -1 if malformed
 * (underflow, or final depth != 1). */
int sro_stack_depth(const int *tok, int len);

/* ---- subtree hashing / detection / cache marking ---- */
int sro_subtree_size(const int *tok, int start, int total); /* detect.cpp:23-47 */
uint64_t sro_hash_flat(const int *tok, const double *val, int len); /* detect.cpp:50-68 */
/* Merkle hashes + sizes for every position (detect.cpp:97-118,161-203). */
void sro_subtree_merkle(const int *tok, const double *val, int len, uint64_t *hash_out,
                        int *size_out);

typedef struct {
    int num_subs;
    int *sub_expr;      /* [num_subs] representative occurrence (expr)  */
    int *sub_start;     /* [num_subs] representative occurrence (start) */
    int *sub_size;      /* [num_subs] */
    uint64_t *sub_hash; /* [num_subs] Merkle hash */
    int *occ_offset;    /* [num_subs+1] into occ_expr/occ_start */
    int *occ_expr;
    int *occ_start;
    int *hints;         /* same packing as the population: -1, -2 or sub id */
    int64_t covered_tokens;
    double avg_subtree_size;
    double coverage;
} sro_detect_result;

/* detect_common_subtrees (detect.cpp:120-387).  Ties in both sorts are broken
 * deterministically (hash, expr, start) and (size desc, hash asc). */
void sro_detect(const int *tok, const double *val, const int *off, const int *len, int E,
                int min_size, int min_freq, sro_detect_result *res);
void sro_detect_free(sro_detect_result *res);

/* Persistent cache keyed by flat hash (cl_gpu_eval.cu:384-397), at most
 * `capacity` slots.  The reference keys by the flat hash alone; that hash
 * keeps a difference confined to the high bits of a value in the high bits
 * of the hash, so subtrees such as x0*x0 and x2*x2 (values 0.0 and 2.0)
 * collide and would share a column.  Here each slot also keeps its subtree's
 * tokens and values and a lookup compares them.  sro_cache_update returns the
 * number of new slots and writes the detected-subtree index of each new slot
 * into new_sub[] and the slot into new_slot[]. */
typedef struct sro_cache_keys sro_cache_keys;
sro_cache_keys *sro_cache_keys_create(int capacity);
void sro_cache_keys_free(sro_cache_keys *k);
int sro_cache_keys_get(const sro_cache_keys *k, uint64_t *keys_out); /* returns the slot count */
int sro_cache_update(const sro_detect_result *det, const int *tok, const double *val,
                     const int *off, sro_cache_keys *k, int *new_sub, int *new_slot);
/* mark_cached_subtrees_from_state (cl_gpu_eval.cu:294-329): left-to-right scan,
 * first cached subtree of size >= min_size wins.  hints packed like tok. */
void sro_mark_cached(const int *tok, const double *val, const int *off, const int *len,
                     int E, const sro_cache_keys *k, int min_size, int *hints);
/* Skeleton build (cl_gpu_eval.cu:433-459): cached subtree -> (OP_REF, slot).
 * Output arrays sized >= total tokens; returns total skeleton tokens. */
int sro_skeleton(const int *tok, const double *val, const int *off, const int *len, int E,
                 const int *hints, int *sk_tok, double *sk_val, int *sk_off, int *sk_len);
/* Evaluate a skeleton point: OP_REF pushes cache[slot*N + dp]. */
void sro_predict32_cached(const int *tok, const double *val, int len, const float *X, int nv,
                          int64_t N, const float *cache, float *out);

/* ---- CPU subtree evaluator, FP64 (cpu_subtree.cpp:83-501) ----
 * X: column-major doubles [(nv+1) * N] (y in column nv, as the reference's
 * per-expression vars[0..nv]); pred [E * N].  threads <= 0 -> 1; results do
 * not depend on the thread count.
 * sro_subtree_predict64 restates eval_cpu_common_subtree_batch
 * (cpu_subtree.cpp:135-340): sro_detect with min_size 3 and min_freq 5 (2
 * below 10 expressions), one FP64 column per detected subtree, then every
 * expression evaluated right to left with the hint map (-2 skip, h >= 0 push
 * column h).  Returns the number of detected subtrees; hints_out (nullable)
 * receives the hint map packed like tok. */
int sro_subtree_predict64(const int *tok, const double *val, const int *off, const int *len, int E,
                          const double *X, int nv, int64_t N, int threads, double *pred, int *hints_out);
/* The stateful form (eval_cpu_stateful_batch, cpu_subtree.cpp:445-501 with
 * detect_and_update_cache, detect.cpp:391-560): subtrees of size >= 2 are
 * grouped by the flat hash of their own tokens; a group already cached is a
 * hit, a new group occurring >= min_freq times (5, or 2 below 10 expressions)
 * joins the cache (ids in ascending hash order); groups are marked largest
 * first, an occurrence overlapping a marked span is skipped; new columns are
 * filled, then every expression is evaluated with the hints.  The reference
 * hashes the first `size` tokens of the EXPRESSION (compute_hash(tokens[i],
 * values[i], size), detect.cpp:429-431) instead of the subtree at j, so its
 * groups mix different subtrees and its predictions are wrong; this
 * restatement keys each subtree by its own hash.  A state belongs to one
 * dataset: a call with another N clears it.  Returns the number of new cache
 * entries. */
typedef struct sro_subtree_state sro_subtree_state;
sro_subtree_state *sro_subtree_state_create(void);
void sro_subtree_state_free(sro_subtree_state *s);
int sro_subtree_state_size(const sro_subtree_state *s);
int sro_subtree_state_predict64(sro_subtree_state *s, const int *tok, const double *val, const int *off,
                                const int *len, int E, const double *X, int nv, int64_t N, int threads,
                                double *pred, int *hints_out);

/* ---- counter-based RNG ---- */
void sro_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* GA RNG hooks.  Default binding is the Philox stream below; the reference
 * harness rebinds them to std::mt19937 to pin the algorithm. */
typedef struct sro_rng {
    float (*uniform)(struct sro_rng *r, float lo, float hi);
    int (*randint)(struct sro_rng *r, int lo, int hi); /* inclusive, lo if hi<=lo */
    void *state;
} sro_rng;

typedef struct {
    uint32_t key[2];
    uint32_t c1, c2, c3; /* counter words 1..3 identify the stream */
    uint32_t word;       /* next 32-bit word of the stream */
} sro_philox_stream;

void sro_philox_stream_init(sro_philox_stream *s, uint64_t seed, uint32_t stream,
                            uint32_t gen, uint32_t domain);
uint32_t sro_philox_next(sro_philox_stream *s);
void sro_rng_bind_philox(sro_rng *r, sro_philox_stream *s);

typedef struct {
    int num_vars;
    int max_tokens;
    int max_depth;
    float const_min, const_max;
    float prob_leaf, prob_const_leaf, prob_unary;
    int n_unary, n_binary;
    int unary_ops[32];
    int binary_ops[32];
} sro_ga_params;

/* Reference defaults (utils.h:178-187, utils.cpp:938-947). */
void sro_ga_params_default(sro_ga_params *p);
int sro_op_arity(int op); /* opcodes.h:42-78, with opcode 29 unary */

/* generate_random_expression (utils.cpp:1018-1161).  Returns length. */
int sro_ga_generate(const sro_ga_params *p, int max_tokens, sro_rng *r, int *tok, float *val);
int sro_ga_subtree_size(const int *tok, int len, int start); /* utils.cpp:1110-1129 */
/* mutate_expression (utils.cpp:1163-1204). */
int sro_ga_mutate(const sro_ga_params *p, const int *ptok, const float *pval, int plen,
                  sro_rng *r, int *ctok, float *cval);
/* crossover_expressions (utils.cpp:1206-1272). */
int sro_ga_crossover(const sro_ga_params *p, const int *ltok, const float *lval, int llen,
                     const int *rtok, const float *rval, int rlen, sro_rng *r, int *ctok,
                     float *cval);

typedef struct {
    int tournament;    /* reference: 3 */
    float p_mutation;  /* reference: 0.2 */
    double elite_frac; /* reference: 0.05 */
} sro_select_params;

/* Sort-by-fitness order (NaN last, ties by index): order[rank] = index. */
void sro_fitness_order(const float *fit, int n, int *order);
/* One generation of selection + variation (cl_sr_system.cpp:285-330).
 * Population stored with fixed stride cap: tok[i*cap + k].  Child c uses the
 * Philox stream (seed, c, gen, domain 1).  Returns 0 or -1 on overflow. */
int sro_ga_next_generation(const sro_ga_params *p, const sro_select_params *sp, int pop,
                           int cap, const int *tok, const float *val, const int *len,
                           const float *fit, uint64_t seed, uint32_t gen, int *out_tok,
                           float *out_val, int *out_len);
/* The same loop with the random stream of child c supplied by rng_for(ud, c)
 * (so a test can script every draw); parent1 / parent2 (nullable, [pop])
 * receive each child's parents: an elite gets its index and -2, a mutant -1
 * as parent2. */
int sro_ga_next_generation_hooked(const sro_ga_params *p, const sro_select_params *sp, int pop, int cap,
                                  const int *tok, const float *val, const int *len, const float *fit,
                                  sro_rng *(*rng_for)(void *ud, int child), void *ud, int *out_tok,
                                  float *out_val, int *out_len, int *parent1, int *parent2);
/* Initial population: individual i uses stream (seed, i, 0xFFFFFFFF, domain 2). */
int sro_ga_init_population(const sro_ga_params *p, int pop, int cap, uint64_t seed,
                           int *out_tok, float *out_val, int *out_len);

/* evogp layout (defs.h:30-107): per individual max_len slots of value,
 * node type (VAR 0, CONST 1, UFUNC 2, BFUNC 3; bit 7 ignored) and subtree
 * size.  To prefix: gpu_simple_jinha_with_evolve.cu:846-866 (a function
 * node's value is its opcode); size[0] outside 1..min(max_len, cap) -> length
 * 0.  From prefix: types from token arity, sizes by brute-force subtree
 * extent; an over-long or malformed stream -> size[0] = 0. */
void sro_evogp_to_prefix(int pop, int max_len, const float *value, const short *type, const short *size,
                         int cap, int *tok, float *val, int *len);
void sro_prefix_to_evogp(int pop, int cap, const int *tok, const float *val, const int *len, int max_len,
                         float *value, short *type, short *size);

#ifdef __cplusplus
}
#endif
#endif
