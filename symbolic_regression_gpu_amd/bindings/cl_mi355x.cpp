// cl_mi355x.cpp -- the reference's GPU evaluation wrappers on top of
// libsrhip.so, for a maintainer to add to the reference build.
//
// The reference declares these functions in src/cleaned/cl_sr_system.cpp:15-28
// and calls them at :206-216 (compile-time choice USE_GPU_SUBTREE /
// USE_GPU_SIMPLE) with an InputInfo (utils.h) and float*** all_vars, where
// all_vars[i][v][dp] is the shared X|y (y at v = num_vars).  Built against
// the reference's src/utils/utils.h (include path: the reference's src/) and
// include/sr_hip.h of this repository; link with -lsrhip.  Errors follow the
// reference's policy for its GPU paths: print and exit.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "sr_hip.h"
#include "utils/utils.h"

namespace {

struct Mi355xCtx {
    sr_ctx *ctx = nullptr;
    int mode = SR_EVAL_BATCHED;
};

void check(sr_ctx *c, int rc, const char *what)
{
    if (rc != SR_OK) {
        std::fprintf(stderr, "%s: %s\n", what, c ? sr_last_error(c) : "no context");
        std::exit(EXIT_FAILURE);
    }
}

void evaluate(InputInfo &in, float ***all_vars, std::vector<float> &mses, Mi355xCtx *m, bool upload_X,
              bool clear_cache, RunStats &rs)
{
    const int E = in.num_exprs, nv = in.num_vars[0];
    const long long N = in.num_dps[0];
    if (upload_X) {  // shared X|y -> column-major [nv + 1][N]
        std::vector<float> X((size_t)(nv + 1) * N);
        for (int v = 0; v <= nv; ++v) std::copy(all_vars[0][v], all_vars[0][v] + N, X.begin() + (size_t)v * N);
        check(m->ctx, sr_set_data(m->ctx, X.data(), nv, N), "sr_set_data");
    }
    if (clear_cache) check(m->ctx, sr_cache_clear(m->ctx), "sr_cache_clear");
    std::vector<int32_t> tok, off(E), len(E);
    std::vector<double> val;
    for (int e = 0; e < E; ++e) {
        off[e] = (int32_t)tok.size();
        len[e] = in.num_tokens[e];
        tok.insert(tok.end(), in.tokens[e], in.tokens[e] + len[e]);
        val.insert(val.end(), in.values[e], in.values[e] + len[e]);
    }
    sr_population pop{E, tok.data(), val.data(), nullptr, off.data(), len.data()};
    sr_stats st{};
    mses.resize(E);
    check(m->ctx, sr_eval_mse(m->ctx, &pop, m->mode, mses.data(), nullptr, &st), "sr_eval_mse");
    rs.total_eval_time_ms = st.total_eval_time_ms;
    rs.drift_detect_time_ms = st.detect_time_ms;
    rs.data_transfer_time_ms = st.transfer_time_ms;
    rs.gpu_kernel_time_ms = st.kernel_time_ms;
    rs.jit_compile_time_ms = st.jit_compile_time_ms;
    rs.num_subtrees = st.num_subtrees;
    rs.avg_subtree_size = st.avg_subtree_size;
    rs.coverage = st.coverage;
}

void *create(int mode)
{
    auto *m = new Mi355xCtx;
    m->mode = mode;
    if (sr_ctx_create(0, &m->ctx) != SR_OK) {
        std::fprintf(stderr, "sr_ctx_create failed: no usable GPU\n");
        std::exit(EXIT_FAILURE);
    }
    // The reference's cache policy is the default; a GA loop gains from the
    // generational one (same MSE bits, columns follow the population).
    if (mode == SR_EVAL_CACHED && std::getenv("SR_CACHE_GENERATIONAL"))
        check(m->ctx, sr_cache_policy(m->ctx, SR_CACHE_GENERATIONAL_SHARED, 0, 3), "sr_cache_policy");
    return m;
}

void destroy(void *p)
{
    auto *m = static_cast<Mi355xCtx *>(p);
    sr_ctx_destroy(m->ctx);
    delete m;
}

}  // namespace

// cl_gpu_eval.cu:569-592 (subtree-cached path)
void *create_gpu_context() { return create(SR_EVAL_CACHED); }
void destroy_gpu_context(void *c) { destroy(c); }
void evaluate_gpu_mse_wrapper(InputInfo &in, float ***X, std::vector<float> &mses, void *c, bool upload_X,
                              bool clear_cache, RunStats &rs)
{
    evaluate(in, X, mses, static_cast<Mi355xCtx *>(c), upload_X, clear_cache, rs);
}

// cl_gpu_simple.cu:315-336 (plain path; the batched launch is the default here)
void *create_gpu_simple_context() { return create(SR_EVAL_BATCHED); }
void destroy_gpu_simple_context(void *c) { destroy(c); }
void evaluate_gpu_simple_wrapper(InputInfo &in, float ***X, std::vector<float> &mses, void *c, bool upload_X,
                                 RunStats &rs)
{
    evaluate(in, X, mses, static_cast<Mi355xCtx *>(c), upload_X, false, rs);
}
