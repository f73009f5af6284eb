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
// result.  The skeleton launches of the device-planned subtree cache
// (cache_kernel, cache_dpt) are tuned separately by sr_autotune_cache, on
// the plan of the last cached generation: both of its interpreter launches
// are re-run with every candidate (they recompute identical columns and
// partial sums) and timed.
//
// Entries are keyed by (architecture, CU count, variable-register class,
// datapoint count) and can be written to and read from a text file
// (sr_tune_save / sr_tune_load), one line per entry:
//   arch cus xc N dpt chunk n_queues blocks_per_cu cache_dpt kernel cache_kernel
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <tuple>

#include "sr_ctx.h"
#include "sr_kernels.h"
#include "sr_launch.h"

namespace sr {

int sr_param_time(Ctx &c, bool adjoint, int dpt, const float *d_cvals, float *ms);  // sr_param_host.cpp

using TuneKey = std::tuple<std::string, int, int, int64_t>;  // arch, CUs, variable class, N

static std::map<TuneKey, sr_tuning> &tune_table()
{
    static std::map<TuneKey, sr_tuning> t;
    return t;
}

static std::mutex &tune_mu()
{
    static std::mutex m;
    return m;
}

static TuneKey tune_key(const Ctx &c) { return TuneKey(c.arch, c.num_cu, pick_xc(c.nv), c.N); }

static bool tuning_valid(const sr_tuning &t)
{
    return t.dpt >= 1 && t.dpt <= 4 && t.chunk >= 64 && (t.n_queues == 1 || t.n_queues == 8) &&
           t.blocks_per_cu >= 0 && t.cache_dpt >= 0 && t.cache_dpt <= 4 && (t.kernel == 0 || t.kernel == 1) &&
           (t.cache_kernel == 0 || t.cache_kernel == 1);
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
    const TuneKey key = tune_key(c);
    {
        std::lock_guard<std::mutex> tl(tune_mu());
        auto it = tune_table().find(key);
        if (it != tune_table().end()) {
            c.tune = it->second;
            if (best_out) *best_out = c.tune;
            if (best_ms) *best_ms = 0.0;
            return SR_OK;
        }
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
    {
        std::lock_guard<std::mutex> tl(tune_mu());
        tune_table()[key] = best;
    }
    if (best_out) *best_out = best;
    if (best_ms) *best_ms = best_t;
    return SR_OK;
}

int sr_autotune_cache(sr_ctx *x, sr_tuning *best_out, double *best_ms)
{
    if (!x) return SR_ERR_INVALID_ARG;
    Ctx &c = x->c;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    hipSetDevice(c.device);
    if (!c.memo.device || !c.dm.ready || !c.dm.planned || c.dm.data_version != c.data_version)
        return c.fail(SR_ERR_INVALID_ARG,
                      "sr_autotune_cache: evaluate a generation with the device-planned cache (SR_CACHE_DEVICE) "
                      "on the current data first");
    const sr_tuning saved = c.tune;
    sr_tuning best = saved;
    double best_t = 1e300;
    hipEvent_t ev[4] = {c.ev[4], c.ev[5], c.ev2[4], c.ev2[5]};
    std::vector<std::tuple<int, int, int>> seen;  // (kind, dpt, xc) of the plans already timed
    // dpt 0 stands for the vector interpreter
    for (int dpt : {0, 1, 2, 3, 4}) {
        c.tune = saved;
        c.tune.cache_kernel = dpt == 0 ? KERNEL_VEC : KERNEL_LANE;
        if (dpt > 0) c.tune.cache_dpt = dpt;
        LaunchPlan lp;
        int rc = plan_batched(c, MODE_MSE_REF, std::max(c.dm.L1, c.dm.E), lp);
        if (rc == SR_ERR_UNSUPPORTED) continue;
        if (rc) {
            c.tune = saved;
            return rc;
        }
        const auto id = std::make_tuple(lp.kind, lp.kind == KERNEL_VEC ? 0 : lp.dpt, lp.xc);
        if (std::find(seen.begin(), seen.end(), id) != seen.end()) continue;
        seen.push_back(id);
        double t = 0.0;
        for (int rep = 0; rep < 3; ++rep) {  // the first run warms caches
            rc = dmemo_launch(c, c.stream, ev);
            hipError_t he = rc ? hipSuccess : hipEventSynchronize(ev[3]);
            if (rc || he != hipSuccess) {
                c.tune = saved;
                return rc ? rc : c.hip_fail(he, "skeleton tuning");
            }
            float t1 = 0.f, t2 = 0.f;
            hipEventElapsedTime(&t1, ev[0], ev[1]);
            hipEventElapsedTime(&t2, ev[2], ev[3]);
            if (rep > 0) t += t1 + t2;
        }
        if (t < best_t) {
            best_t = t;
            best = c.tune;
        }
    }
    c.tune = best;
    c.ga.release_graphs();  // captured generations bake the skeleton kernel in
    {
        std::lock_guard<std::mutex> tl(tune_mu());
        auto it = tune_table().find(tune_key(c));
        if (it != tune_table().end()) {
            it->second.cache_kernel = best.cache_kernel;
            it->second.cache_dpt = best.cache_dpt;
        } else {
            tune_table()[tune_key(c)] = best;
        }
    }
    if (best_out) *best_out = best;
    if (best_ms) *best_ms = best_t / 2;
    return SR_OK;
}

int sr_tune_save(const char *path)
{
    if (!path) return SR_ERR_INVALID_ARG;
    FILE *f = std::fopen(path, "w");
    if (!f) return SR_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> tl(tune_mu());
    std::fprintf(f, "# arch cus xc N dpt chunk n_queues blocks_per_cu cache_dpt kernel cache_kernel\n");
    for (const auto &kv : tune_table()) {
        const sr_tuning &t = kv.second;
        std::fprintf(f, "%s %d %d %lld %d %d %d %d %d %d %d\n", std::get<0>(kv.first).c_str(), std::get<1>(kv.first),
                     std::get<2>(kv.first), (long long)std::get<3>(kv.first), t.dpt, t.chunk, t.n_queues,
                     t.blocks_per_cu, t.cache_dpt, t.kernel, t.cache_kernel);
    }
    return std::fclose(f) == 0 ? SR_OK : SR_ERR_INVALID_ARG;
}

int sr_tune_load(const char *path, int32_t *entries)
{
    if (!path) return SR_ERR_INVALID_ARG;
    FILE *f = std::fopen(path, "r");
    if (!f) return SR_ERR_INVALID_ARG;
    std::vector<std::pair<TuneKey, sr_tuning>> got;
    char line[512];
    int rc = SR_OK;
    while (std::fgets(line, sizeof line, f)) {
        if (line[0] == '#' || line[0] == '\n') continue;
        char arch[64];
        int cus = 0, xc = 0;
        long long n = 0;
        sr_tuning t{};
        if (std::sscanf(line, "%63s %d %d %lld %d %d %d %d %d %d %d", arch, &cus, &xc, &n, &t.dpt, &t.chunk,
                        &t.n_queues, &t.blocks_per_cu, &t.cache_dpt, &t.kernel, &t.cache_kernel) != 11 ||
            cus <= 0 || (xc != 4 && xc != 16 && xc != 32) || n <= 0 || !tuning_valid(t)) {
            rc = SR_ERR_INVALID_ARG;
            break;
        }
        got.emplace_back(TuneKey(arch, cus, xc, (int64_t)n), t);
    }
    std::fclose(f);
    if (rc) return rc;  // nothing is taken from a file with a bad line
    std::lock_guard<std::mutex> tl(tune_mu());
    for (auto &kv : got) tune_table()[kv.first] = kv.second;
    if (entries) *entries = (int32_t)got.size();
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
