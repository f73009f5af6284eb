// sr_tune.cpp -- autotuner for the launch shape of the batched interpreter
// and of the constant-fitting kernels.
//
// The batched launch has these knobs (sr_tuning): the interpreter (lane:
// variables and stack in VGPRs, DPT 1..4 datapoints per lane -- dispatch
// amortised over more datapoints vs. more registers per wave, fewer waves
// per SIMD; or vector: 384-datapoint tiles, stack in LDS, any variable
// count), datapoints per work item (chunk, multiples of 768: the number of
// waves that share one expression, LPT granularity vs. per-item program
// staging), XCD queues (1 or 8; tuned runs keep 8) and persistent workgroups
// per CU (0 = occupancy limit).  The tuner times every valid combination on a
// length-stratified sample of the population (the interpreter kernel alone,
// HIP events), keeps the fastest, and remembers it per (variable-register
// class, datapoint count), so later calls with a similar shape reuse the
// result.  The skeleton launches of the subtree caches (cache_kernel,
// cache_dpt) are left as set.
#include <algorithm>
#include <cstring>
#include <map>
#include <numeric>
#include <tuple>

#include "sr_ctx.h"
#include "sr_kernels.h"
#include "sr_launch.h"

namespace sr {

int sr_param_time(Ctx &c, bool adjoint, int dpt, const float *d_cvals, float *ms);  // sr_param_host.cpp

static std::map<std::tuple<int, int, int64_t>, sr_tuning> &tune_table()
{
    static std::map<std::tuple<int, int, int64_t>, sr_tuning> t;
    return t;
}

}  // namespace sr

using namespace sr;

extern "C" {

int sr_autotune(sr_ctx *x, const sr_population *pop, int32_t max_sample, sr_tuning *best_out, double *best_ms)
{
    if (!x || !pop || pop->num_exprs <= 0) return SR_ERR_INVALID_ARG;
    Ctx &c = x->c;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    hipSetDevice(c.device);
    if (!c.X) return c.fail(SR_ERR_NO_DATA, "no data: call sr_set_data first");
    const int xc = pick_xc(c.nv);
    const auto key = std::make_tuple(c.device, xc, c.N);
    auto it = tune_table().find(key);
    if (it != tune_table().end()) {
        c.tune = it->second;
        if (best_out) *best_out = c.tune;
        if (best_ms) *best_ms = 0.0;
        return SR_OK;
    }
    // length-stratified sample: every k-th expression in length order
    const int E = pop->num_exprs;
    const int S = std::max(1, std::min(E, max_sample > 0 ? max_sample : 256));
    std::vector<int> idx(E);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return pop->lengths[a] > pop->lengths[b]; });
    std::vector<int32_t> off(S), len(S);
    std::vector<int32_t> tok;
    std::vector<double> val;
    std::vector<float> vf;
    for (int k = 0; k < S; ++k) {
        const int e = idx[(int64_t)k * E / S];
        off[k] = (int32_t)tok.size();
        len[k] = pop->lengths[e];
        tok.insert(tok.end(), pop->tokens + pop->offsets[e], pop->tokens + pop->offsets[e] + len[k]);
        expr_values_f32(*pop, e, vf);
        for (float f : vf) val.push_back((double)f);
    }
    sr_population sp{S, tok.data(), val.data(), nullptr, off.data(), len.data()};
    Program p;
    int rc = build_program(c, sp, false, p);
    if (rc) return rc;
    const sr_tuning saved = c.tune;
    sr_tuning best = saved;
    double best_t = 1e300;
    const int chunks[] = {3072, 6144, 12288, 24576};
    // one queue per XCD keeps each chunk's X in one L2; a small sample
    // cannot tell the queue layouts apart, so it is not a tuned knob
    const int queues[] = {8};
    const int bpcs[] = {0, 1};
    // dpt 0 stands for the vector interpreter
    for (int dpt : {0, 1, 2, 3, 4}) {
        if (dpt > 0 && ((xc == 32 && dpt > 2) || (dpt == 3 && xc != 4))) continue;
        for (int ch : chunks) {
            for (int nq : queues)
                for (int bpc : bpcs) {
                    c.tune = sr_tuning{dpt > 0 ? dpt : saved.dpt, ch, nq, bpc, saved.cache_dpt,
                                       dpt > 0 ? KERNEL_LANE : KERNEL_VEC, saved.cache_kernel};
                    double t = 0.0;
                    for (int rep = 0; rep < 2; ++rep) {  // first run warms caches
                        sr_stats st{};
                        rc = run_batched(c, p, MODE_MSE, nullptr, nullptr, nullptr, &st);
                        if (rc) {
                            c.tune = saved;
                            return rc;
                        }
                        t = st.interp_kernel_ms;
                    }
                    if (t < best_t) {
                        best_t = t;
                        best = c.tune;
                    }
                }
        }
    }
    c.tune = best;
    tune_table()[key] = best;
    if (best_out) *best_out = best;
    if (best_ms) *best_ms = best_t;
    return SR_OK;
}

int sr_param_autotune(sr_ctx *x, const float *d_cvals, int32_t *fwd_dpt, int32_t *adj_dpt)
{
    if (!x || !d_cvals) return SR_ERR_INVALID_ARG;
    Ctx &c = x->c;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    hipSetDevice(c.device);
    int best[2] = {c.par.fwd_dpt, c.par.adj_dpt};
    for (int kind = 0; kind < 2; ++kind) {
        float best_t = 1e30f;
        for (int dpt : {1, 2, 4}) {
            float ms = 0.f;
            int rc = sr_param_time(c, kind == 1, dpt, d_cvals, &ms);
            if (rc == SR_ERR_UNSUPPORTED) continue;  // no variant for this shape
            if (rc) return rc;
            if (ms < best_t) {
                best_t = ms;
                best[kind] = dpt;
            }
        }
    }
    c.par.fwd_dpt = best[0];
    c.par.adj_dpt = best[1];
    if (fwd_dpt) *fwd_dpt = best[0];
    if (adj_dpt) *adj_dpt = best[1];
    return SR_OK;
}

}  // extern "C"
