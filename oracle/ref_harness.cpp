// ref_harness.cpp -- C entry points around the reference's own CPU sources,
// compiled where they lie (oracle/Makefile adds -I$(REF)/src).  Output goes to
// oracle/_ref/libref.so only.  Used to pin the oracle restatement and to write
// the golden fixtures under tests/golden/.  Test infrastructure only.
//
// Translation units pulled in:
//   cleaned/cl_cpu_eval.cpp   eval_op_cpu / evaluate_single_expr / evaluate_cpu_mse
//   utils/detect.cpp          compute_hash / hash_node / detect_common_subtrees
//   eval/cpu_simple_single.cpp eval_op (fp64 twin)           [linked separately]
//   utils/utils.cpp           GA operators + set_seed        [linked separately]
//   eval/cpu_subtree.cpp      eval_cpu_common_subtree_batch / eval_cpu_stateful_batch
//                             (built with -DUSE_CPU_SUBTREE) [linked separately]
#include "cleaned/cl_cpu_eval.cpp"
#include "utils/detect.cpp"
#define USE_CPU_SUBTREE
#include "eval/evaluator.h"

#include <cstdint>
#include <random>
#include <vector>

#include "sr_oracle.h"

double eval_op(int op, double val1, double val2);  // cpu_simple_single.cpp:6

extern "C" {

float ref_op32(int op, float a, float b) { return eval_op_cpu(op, a, b); }
double ref_op64(int op, double a, double b) { return eval_op(op, a, b); }

float ref_eval_point32(const int *tok, const double *val, int len, const float *x, int nv)
{
    return evaluate_single_expr(tok, val, len, x, nv);
}

// evaluate_cpu_mse over a packed population sharing one column-major X|y.
void ref_mse32(const int *tok, const double *val, const int *off, const int *len, int E,
               const float *X, int nv, int64_t N, float *mse_out)
{
    InputInfo info{};
    info.num_exprs = E;
    std::vector<int> nvars(E, nv), ndps(E, (int)N), ntok(E);
    std::vector<int *> toks(E);
    std::vector<double *> vals(E);
    for (int e = 0; e < E; e++) {
        ntok[e] = len[e];
        toks[e] = const_cast<int *>(tok + off[e]);
        vals[e] = const_cast<double *>(val + off[e]);
    }
    info.num_vars = nvars.data();
    info.num_dps = ndps.data();
    info.num_tokens = ntok.data();
    info.tokens = toks.data();
    info.values = vals.data();
    std::vector<float *> cols(nv + 1);
    for (int v = 0; v <= nv; v++) cols[v] = const_cast<float *>(X + (int64_t)v * N);
    std::vector<float **> all(E, cols.data());
    std::vector<float> mses;
    RunStats st;
    evaluate_cpu_mse(info, all.data(), mses, st);
    for (int e = 0; e < E; e++) mse_out[e] = mses[e];
}

uint64_t ref_hash_flat(const int *tok, const double *val, int len)
{
    return compute_hash(tok, val, len);
}

uint64_t ref_hash_node(int token, double value, const uint64_t *children, int nch)
{
    std::vector<uint64_t> ch(children, children + nch);
    return hash_node(token, value, ch);
}

// detect_common_subtrees: returns num_subs; fills per-sub size/occ count, the
// flat hash of each detected subtree, and the hint map (packed like tok).
int ref_detect(const int *tok, const double *val, const int *off, const int *len, int E,
               int min_size, int min_freq, int *sub_size, int *sub_occ, uint64_t *sub_flat,
               int *hints, int max_subs)
{
    std::vector<int> ntok(len, len + E);
    std::vector<const int *> toks(E);
    std::vector<const double *> vals(E);
    for (int e = 0; e < E; e++) {
        toks[e] = tok + off[e];
        vals[e] = val + off[e];
    }
    SubtreeDetectionResult r = detect_common_subtrees(E, 1, ntok.data(), toks.data(),
                                                      vals.data(), min_size, min_freq);
    int n = r.num_subs;
    for (int s = 0; s < n && s < max_subs; s++) {
        sub_size[s] = r.num_sub_tokens[s];
        sub_occ[s] = r.num_occurrences[s];
        sub_flat[s] = compute_hash(r.sub_tokens[s], r.sub_values[s], r.num_sub_tokens[s]);
    }
    for (int e = 0; e < E; e++)
        for (int i = 0; i < len[e]; i++) hints[off[e] + i] = r.expr_sub_hints[e][i];
    free_subtree_detection_result(r);
    return n;
}

// ---- GA: reference operators vs the oracle restatement under one mt19937 ----

static std::mt19937 g_twin;

static float twin_uniform(sro_rng *, float lo, float hi)
{
    std::uniform_real_distribution<float> d(lo, hi);
    return d(g_twin);
}

static int twin_randint(sro_rng *, int lo, int hi)
{
    if (hi <= lo) return lo;
    std::uniform_int_distribution<int> d(lo, hi);
    return d(g_twin);
}

static EvolutionParams to_ref(const sro_ga_params *p)
{
    EvolutionParams e;
    e.num_vars = p->num_vars;
    e.max_tokens = p->max_tokens;
    e.max_depth = p->max_depth;
    e.const_min = p->const_min;
    e.const_max = p->const_max;
    e.prob_leaf = p->prob_leaf;
    e.prob_const_leaf = p->prob_const_leaf;
    e.prob_unary = p->prob_unary;
    return e;
}

static int out(const std::vector<int> &t, const std::vector<float> &v, int *tok, float *val)
{
    for (size_t i = 0; i < t.size(); i++) {
        tok[i] = t[i];
        val[i] = v[i];
    }
    return (int)t.size();
}

// Draw log of the oracle run under the mt19937 twin (kind 0 = uniform(lo, hi)
// with the float result's bits, kind 1 = randint(lo, hi)).
static int *g_log_kind, *g_log_res, g_log_n, g_log_max;
static float *g_log_lo, *g_log_hi;

static void log_draw(int kind, float lo, float hi, int res)
{
    if (g_log_n < g_log_max) {
        g_log_kind[g_log_n] = kind;
        g_log_lo[g_log_n] = lo;
        g_log_hi[g_log_n] = hi;
        g_log_res[g_log_n] = res;
    }
    g_log_n++;
}

static float logged_uniform(sro_rng *r, float lo, float hi)
{
    float f = twin_uniform(r, lo, hi);
    int bits;
    memcpy(&bits, &f, 4);
    log_draw(0, lo, hi, bits);
    return f;
}

static int logged_randint(sro_rng *r, int lo, int hi)
{
    int v = twin_randint(r, lo, hi);
    log_draw(1, (float)lo, (float)hi, v);
    return v;
}

// Oracle GA under the twin RNG, recording every draw.  Returns the output
// length; *n_log = number of draws (may exceed max_log: then truncated).
int ref_ga_op_log(int op, int seed, const sro_ga_params *p, const int *at, const float *av, int alen,
                  const int *bt, const float *bv, int blen, int *tok, float *val, int *kind,
                  float *lo, float *hi, int *res, int max_log, int *n_log)
{
    g_twin.seed(seed);
    g_log_kind = kind;
    g_log_lo = lo;
    g_log_hi = hi;
    g_log_res = res;
    g_log_n = 0;
    g_log_max = max_log;
    sro_rng r{logged_uniform, logged_randint, nullptr};
    int n;
    if (op == 0) n = sro_ga_generate(p, p->max_tokens, &r, tok, val);
    else if (op == 1) n = sro_ga_mutate(p, at, av, alen, &r, tok, val);
    else n = sro_ga_crossover(p, at, av, alen, bt, bv, blen, &r, tok, val);
    *n_log = g_log_n;
    return n;
}

// op: 0 generate, 1 mutate(a), 2 crossover(a, b).  which: 0 reference, 1 oracle.
int ref_ga_op(int op, int which, int seed, const sro_ga_params *p, const int *at,
              const float *av, int alen, const int *bt, const float *bv, int blen, int *tok,
              float *val)
{
    if (which == 0) {
        set_seed(seed);
        EvolutionParams e = to_ref(p);
        std::vector<int> t;
        std::vector<float> v;
        std::vector<int> a(at, at + alen), b(bt, bt + blen);
        std::vector<float> fa(av, av + alen), fb(bv, bv + blen);
        if (op == 0) generate_random_expression(e, t, v);
        else if (op == 1) mutate_expression(e, a, fa, t, v);
        else crossover_expressions(e, a, fa, b, fb, t, v);
        return out(t, v, tok, val);
    }
    g_twin.seed(seed);
    sro_rng r{twin_uniform, twin_randint, nullptr};
    if (op == 0) return sro_ga_generate(p, p->max_tokens, &r, tok, val);
    if (op == 1) return sro_ga_mutate(p, at, av, alen, &r, tok, val);
    return sro_ga_crossover(p, at, av, alen, bt, bv, blen, &r, tok, val);
}

// parse_input_info (utils.cpp:122-233) of a digest file.  Returns the
// expression count (0 on the reference's own failure) and fills the arrays
// when they are large enough: per expression nv / ndps / ntok, packed
// tokens and values, paths at path_stride bytes each, the shared flag.
int ref_parse_digest(const char *path, int cap_exprs, int cap_tokens, int path_stride, int *nv, int *ndps,
                     int *ntok, int *tokens, double *values, char *paths, int *shared)
{
    InputInfo info = parse_input_info(path);
    const int E = info.num_exprs;
    if (E > 0 && E <= cap_exprs) {
        int t = 0;
        for (int e = 0; e < E; ++e) {
            nv[e] = info.num_vars[e];
            ndps[e] = info.num_dps[e];
            ntok[e] = info.num_tokens[e];
            for (int k = 0; k < ntok[e] && t < cap_tokens; ++k, ++t) {
                tokens[t] = info.tokens[e][k];
                values[t] = info.values[e][k];
            }
            std::snprintf(paths + (size_t)e * path_stride, path_stride, "%s", info.data_filenames[e].c_str());
        }
        *shared = info.has_shared_data ? 1 : 0;
    }
    if (E > 0) free_input_info(info);
    return E;
}

// load_data_file (utils.cpp:235-300): out = column-major [nv+1][ndps].
void ref_load_data(const char *path, int nv, int ndps, double *out)
{
    double **v = load_data_file(path, nv, ndps);
    for (int i = 0; i <= nv; ++i) {
        for (int j = 0; j < ndps; ++j) out[(size_t)i * ndps + j] = v[i][j];
        free(v[i]);
    }
    free(v);
}

// ---- CPU subtree evaluator (cpu_subtree.cpp:135-501) on one shared dataset --
// X column-major [(nv+1)][N]; pred [E][N].
struct RefBatch {
    InputInfo info{};
    std::vector<int> nv, nd, nt;
    std::vector<int *> tp;
    std::vector<double *> vp;
    std::vector<double *> cols;
    std::vector<double **> vars;
    std::vector<double *> preds;
};

static void ref_batch(RefBatch &b, const int *tok, const double *val, const int *off, const int *len, int E,
                      const double *X, int nv, int N, double *pred)
{
    b.nv.assign(E, nv);
    b.nd.assign(E, N);
    b.nt.assign(len, len + E);
    b.tp.resize(E);
    b.vp.resize(E);
    b.preds.resize(E);
    b.cols.resize(nv + 1);
    for (int i = 0; i <= nv; ++i) b.cols[i] = const_cast<double *>(X) + (size_t)i * N;
    b.vars.assign(E, b.cols.data());
    int maxt = 0;
    for (int e = 0; e < E; ++e) {
        b.tp[e] = const_cast<int *>(tok + off[e]);
        b.vp[e] = const_cast<double *>(val + off[e]);
        b.preds[e] = pred + (size_t)e * N;
        maxt = std::max(maxt, len[e]);
    }
    b.info.num_exprs = E;
    b.info.num_vars = b.nv.data();
    b.info.num_dps = b.nd.data();
    b.info.num_tokens = b.nt.data();
    b.info.tokens = b.tp.data();
    b.info.values = b.vp.data();
    b.info.max_tokens = maxt;
    b.info.max_num_dps = N;
    b.info.max_num_features = nv + 1;
    b.info.has_shared_data = true;
}

void ref_subtree_predict64(const int *tok, const double *val, const int *off, const int *len, int E,
                           const double *X, int nv, int N, double *pred)
{
    RefBatch b;
    ref_batch(b, tok, val, off, len, E, X, nv, N, pred);
    eval_cpu_common_subtree_batch(b.info, b.vars.data(), b.preds.data(), nullptr);
}

void *ref_subtree_cache_create() { return new SubtreeCache(); }
void ref_subtree_cache_free(void *c) { delete static_cast<SubtreeCache *>(c); }

int ref_subtree_stateful_predict64(void *c, const int *tok, const double *val, const int *off, const int *len,
                                   int E, const double *X, int nv, int N, double *pred)
{
    RefBatch b;
    ref_batch(b, tok, val, off, len, E, X, nv, N, pred);
    SubtreeCache &cache = *static_cast<SubtreeCache *>(c);
    eval_cpu_stateful_batch(b.info, b.vars.data(), b.preds.data(), cache, nullptr);
    return (int)cache.results.size();
}

}  // extern "C"
