// sr_ga_host.cpp -- C ABI of the device-resident GA loop: population storage,
// device evaluation, selection + variation, and the per-generation hipGraph.
//
// One generation in BATCHED mode is, all on the device:
//   k_compile_pop -> k_order_by_len -> zero counters -> k_batched (MSE) ->
//   k_finalize -> k_ga_rank (+ fitness and length history) -> k_ga_next -> gen+1
// (cl_sr_system.cpp:168-333 runs the same sequence with host selection and
// variation).  The two buffer parities are captured as two graphs and replayed
// alternately.
#include <algorithm>
#include <chrono>
#include <cmath>

#include "sr_ctx.h"
#include "sr_ga_launch.h"
#include "sr_kernels.h"
#include "sr_launch.h"

static_assert(sizeof(sr_ga_params) == sizeof(sr::GaParams), "sr_ga_params layout");

namespace sr {

static int ga_ready(Ctx &c)
{
    if (!c.ga.ready) return c.fail(SR_ERR_INVALID_ARG, "no GA population: call sr_ga_init first");
    return SR_OK;
}

static int ga_alloc(Ctx &c, int pop, int cap)
{
    GaState &g = c.ga;
    if (g.pop != pop || g.cap != cap) g.release_graphs();
    const size_t pc = (size_t)pop * cap;
    hipError_t he = hipSuccess;
    for (int b = 0; b < 2 && he == hipSuccess; ++b) {
        if ((he = g.tok[b].ensure(pc * 4)) != hipSuccess) break;
        if ((he = g.val[b].ensure(pc * 4)) != hipSuccess) break;
        he = g.len[b].ensure((size_t)pop * 4);
    }
    if (he == hipSuccess) he = g.progs.ensure(pc * sizeof(Ins));
    if (he == hipSuccess) he = g.prog_len.ensure((size_t)pop * 4);
    if (he == hipSuccess) he = g.size_s.ensure(pc * 4);
    if (he == hipSuccess) he = g.need_s.ensure(pc * 4);
    if (he == hipSuccess) he = g.frames.ensure((size_t)pop * (cap + 1) * sizeof(CompileFrame));
    if (he == hipSuccess) he = g.order_len.ensure((size_t)pop * 4);
    if (he == hipSuccess) he = g.order_fit.ensure((size_t)pop * 4);
    if (he == hipSuccess) he = g.fit.ensure((size_t)pop * 4);
    if (he == hipSuccess) he = g.status.ensure(64);
    if (he == hipSuccess) he = g.max_need.ensure(64);
    if (he == hipSuccess) he = g.gen.ensure(64);
    // the status word is read after every GA step: it starts cleared, not as
    // whatever the allocation held
    if (he == hipSuccess) he = hipMemsetAsync(g.status.p, 0, 4, c.stream);
    if (he != hipSuccess) return c.hip_fail(he, "hipMalloc GA buffers");
    g.pop = pop;
    g.cap = cap;
    return SR_OK;
}

static int ga_status(Ctx &c)
{
    int32_t st = 0;
    hipError_t he = hipMemcpy(&st, c.ga.status.p, 4, hipMemcpyDeviceToHost);
    if (he != hipSuccess) return c.hip_fail(he, "GA status");
    if (st) {
        hipMemset(c.ga.status.p, 0, 4);
        if (st & 4) return c.fail(SR_ERR_BAD_EXPR, "GA population holds a malformed expression");
        if (st & 8) return c.fail(SR_ERR_UNSUPPORTED, "GA expression needs more than 8 stack slots");
        return c.fail(SR_ERR_CAPACITY, "GA individual exceeded its token capacity");
    }
    return SR_OK;
}

// Device evaluation of population buffer b into c.d_mse / c.d_cnt.
static int ga_enqueue_eval(Ctx &c, int b, const LaunchPlan &lp, hipStream_t s)
{
    GaState &g = c.ga;
    hipError_t he = launch_compile_pop(g.tok[b].as<int32_t>(), g.val[b].as<float>(), g.len[b].as<int32_t>(),
                                       g.pop, g.cap, c.nv, g.progs.as<Ins>(), g.prog_len.as<int32_t>(),
                                       g.size_s.as<int32_t>(), g.need_s.as<int32_t>(),
                                       g.frames.as<CompileFrame>(), g.max_need.as<int32_t>(),
                                       g.status.as<int32_t>(), s);
    if (he == hipSuccess) he = launch_order_by_len(g.prog_len.as<int32_t>(), g.pop, g.order_len.as<int32_t>(), s);
    if (he != hipSuccess) return c.hip_fail(he, "GA compile");
    BatchArgs a{};
    a.progs = g.progs.as<Ins>();
    a.prog_off = nullptr;
    a.prog_stride = g.cap;
    a.prog_len = g.prog_len.as<int32_t>();
    a.order = g.order_len.as<int32_t>();
    return enqueue_batched(c, MODE_MSE, lp, a, g.pop, s);
}

static SelectParams to_select(const sr_select_params &sp, int pop)
{
    SelectParams s;
    s.tournament = sp.tournament < 1 ? 1 : sp.tournament;
    s.p_mutation = sp.p_mutation;
    int e = (int)(pop * sp.elite_frac);
    s.elites = std::min(std::max(e, 1), pop);
    return s;
}

// Rank fitness fit (device [pop]) and write generation *gen's offspring of
// buffer b into buffer 1-b.
static int ga_enqueue_next(Ctx &c, int b, const SelectParams &sp, uint64_t seed, const float *fit,
                           hipStream_t s, const GaRecord &rec = GaRecord{})
{
    GaState &g = c.ga;
    hipError_t he = launch_ga_rank(fit, g.pop, g.order_fit.as<int32_t>(), s, rec);
    if (he == hipSuccess)
        he = launch_ga_next(g.p, sp, g.pop, g.cap, g.tok[b].as<int32_t>(), g.val[b].as<float>(),
                            g.len[b].as<int32_t>(), fit, g.order_fit.as<int32_t>(), seed,
                            g.gen.as<uint32_t>(), g.tok[1 - b].as<int32_t>(), g.val[1 - b].as<float>(),
                            g.len[1 - b].as<int32_t>(), g.status.as<int32_t>(), s);
    if (he != hipSuccess) return c.hip_fail(he, "GA next generation");
    return SR_OK;
}

// Host copy of the current population as a packed sr_population.
static int ga_download(Ctx &c, std::vector<int32_t> &tok, std::vector<double> &val,
                       std::vector<int32_t> &off, std::vector<int32_t> &len)
{
    GaState &g = c.ga;
    const size_t pc = (size_t)g.pop * g.cap;
    std::vector<int32_t> t(pc), l(g.pop);
    std::vector<float> v(pc);
    hipError_t he = hipMemcpyAsync(t.data(), g.tok[g.cur].p, pc * 4, hipMemcpyDeviceToHost, c.stream);
    if (he == hipSuccess) he = hipMemcpyAsync(v.data(), g.val[g.cur].p, pc * 4, hipMemcpyDeviceToHost, c.stream);
    if (he == hipSuccess) he = hipMemcpyAsync(l.data(), g.len[g.cur].p, (size_t)g.pop * 4, hipMemcpyDeviceToHost, c.stream);
    if (he == hipSuccess) he = hipStreamSynchronize(c.stream);
    if (he != hipSuccess) return c.hip_fail(he, "D2H population");
    tok.clear();
    val.clear();
    off.resize(g.pop);
    len = l;
    for (int i = 0; i < g.pop; ++i) {
        off[i] = (int32_t)tok.size();
        for (int k = 0; k < l[i]; ++k) {
            tok.push_back(t[(size_t)i * g.cap + k]);
            val.push_back((double)v[(size_t)i * g.cap + k]);  // float -> double, cl_sr_system.cpp:194-198
        }
    }
    return SR_OK;
}

static uint64_t mix(uint64_t h, uint64_t v)
{
    h ^= v + 0x9e3779b97f4a7c15ULL + (h << 6) + (h >> 2);
    return h;
}

}  // namespace sr

using namespace sr;

extern "C" {

int sr_ga_params_default(sr_ga_params *p)
{
    if (!p) return SR_ERR_INVALID_ARG;
    std::memset(p, 0, sizeof *p);
    p->num_vars = 1;
    p->max_tokens = 64;
    p->max_depth = 5;
    p->const_min = -5.0f;
    p->const_max = 5.0f;
    p->prob_leaf = 0.1f;
    p->prob_const_leaf = 0.5f;
    p->prob_unary = 0.35f;
    p->n_unary = 18;
    for (int k = 0; k < 18; ++k) p->unary_ops[k] = 10 + k;
    p->n_binary = 9;
    for (int k = 0; k < 9; ++k) p->binary_ops[k] = 1 + k;
    return SR_OK;
}

// Validate and adopt GA parameters (population storage follows in ga_alloc).
static int ga_configure(Ctx &c, const sr_ga_params *p)
{
    if (p->max_tokens < 1 || p->max_tokens > MAX_EXPR_LEN - 64 || p->max_depth < 0 || p->max_depth > 60 ||
        p->n_unary < 1 || p->n_unary > GA_MAX_OPS || p->n_binary < 1 || p->n_binary > GA_MAX_OPS)
        return c.fail(SR_ERR_INVALID_ARG, "GA parameters out of range");
    for (int k = 0; k < p->n_unary; ++k)
        if (!((p->unary_ops[k] >= 10 && p->unary_ops[k] <= 27) || p->unary_ops[k] == OP_EXP_CLAMPED))
            return c.fail(SR_ERR_INVALID_ARG, "unary op set holds a non-unary opcode");
    for (int k = 0; k < p->n_binary; ++k)
        if (!(p->binary_ops[k] >= 1 && p->binary_ops[k] <= 9))
            return c.fail(SR_ERR_INVALID_ARG, "binary op set holds a non-binary opcode");
    std::memcpy(&c.ga.p, p, sizeof c.ga.p);
    c.ga.release_graphs();
    return SR_OK;
}

int sr_ga_init(sr_ctx *x, const sr_ga_params *p, int32_t pop, uint64_t seed)
{
    if (!x || !p || pop <= 0) return SR_ERR_INVALID_ARG;
    Ctx &c = x->c;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    hipSetDevice(c.device);
    int rc = ga_configure(c, p);
    if (rc) return rc;
    GaState &g = c.ga;
    int cap = p->max_tokens + p->max_depth + 2;
    rc = ga_alloc(c, pop, cap);
    if (rc) return rc;
    g.cur = 0;
    g.ready = true;
    hipMemsetAsync(g.status.p, 0, 4, c.stream);
    hipError_t he = launch_ga_init(g.p, pop, cap, seed, g.tok[0].as<int32_t>(), g.val[0].as<float>(),
                                   g.len[0].as<int32_t>(), g.status.as<int32_t>(), c.stream);
    if (he == hipSuccess) he = hipStreamSynchronize(c.stream);
    if (he != hipSuccess) return c.hip_fail(he, "GA init");
    return ga_status(c);
}

int sr_ga_get_population(sr_ctx *x, int32_t *tok, float *val, int32_t *len, int32_t *pop, int32_t *cap)
{
    if (!x) return SR_ERR_INVALID_ARG;
    Ctx &c = x->c;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    int rc = ga_ready(c);
    if (rc) return rc;
    GaState &g = c.ga;
    if (pop) *pop = g.pop;
    if (cap) *cap = g.cap;
    const size_t pc = (size_t)g.pop * g.cap;
    hipError_t he = hipSuccess;
    if (tok) he = hipMemcpyAsync(tok, g.tok[g.cur].p, pc * 4, hipMemcpyDeviceToHost, c.stream);
    if (he == hipSuccess && val) he = hipMemcpyAsync(val, g.val[g.cur].p, pc * 4, hipMemcpyDeviceToHost, c.stream);
    if (he == hipSuccess && len)
        he = hipMemcpyAsync(len, g.len[g.cur].p, (size_t)g.pop * 4, hipMemcpyDeviceToHost, c.stream);
    if (he == hipSuccess) he = hipStreamSynchronize(c.stream);
    if (he != hipSuccess) return c.hip_fail(he, "D2H population");
    return SR_OK;
}

int sr_ga_set_population(sr_ctx *x, const int32_t *tok, const float *val, const int32_t *len, int32_t pop,
                         int32_t cap)
{
    if (!x || !tok || !val || !len || pop <= 0) return SR_ERR_INVALID_ARG;
    Ctx &c = x->c;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    int rc = ga_ready(c);
    if (rc) return rc;
    if (cap != c.ga.p.max_tokens + c.ga.p.max_depth + 2)
        return c.fail(SR_ERR_INVALID_ARG, "stride differs from max_tokens + max_depth + 2");
    for (int i = 0; i < pop; ++i)
        if (len[i] < 1 || len[i] > cap) return c.fail(SR_ERR_INVALID_ARG, "individual length outside 1..cap");
    rc = ga_alloc(c, pop, cap);
    if (rc) return rc;
    GaState &g = c.ga;
    const size_t pc = (size_t)pop * cap;
    hipError_t he = hipMemcpyAsync(g.tok[g.cur].p, tok, pc * 4, hipMemcpyHostToDevice, c.stream);
    if (he == hipSuccess) he = hipMemcpyAsync(g.val[g.cur].p, val, pc * 4, hipMemcpyHostToDevice, c.stream);
    if (he == hipSuccess) he = hipMemcpyAsync(g.len[g.cur].p, len, (size_t)pop * 4, hipMemcpyHostToDevice, c.stream);
    if (he == hipSuccess) he = hipStreamSynchronize(c.stream);
    if (he != hipSuccess) return c.hip_fail(he, "H2D population");
    return SR_OK;
}

int sr_ga_evaluate(sr_ctx *x, int mode, float *mse_out, sr_stats *st)
{
    if (!x) return SR_ERR_INVALID_ARG;
    Ctx &c = x->c;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    hipSetDevice(c.device);
    int rc = ga_ready(c);
    if (rc) return rc;
    if (!c.X) return c.fail(SR_ERR_NO_DATA, "no data: call sr_set_data first");
    GaState &g = c.ga;
    g.last_mse.resize(g.pop);
    const bool device_cache = mode == SR_EVAL_CACHED && c.memo.device;
    if (mode == SR_EVAL_BATCHED || device_cache) {
        LaunchPlan lp;
        rc = plan_batched(c, MODE_MSE, g.pop, lp);
        if (rc) return rc;
        long long before[9] = {0};
        if (device_cache) {
            if ((rc = dmemo_setup(c, c.memo.capacity, c.memo.min_size)) != SR_OK) return rc;
            if ((rc = dmemo_stats(c, before)) != SR_OK) return rc;
        }
        auto t0 = std::chrono::steady_clock::now();
        hipEvent_t kev[4] = {c.ev[4], c.ev[5], c.ev2[4], c.ev2[5]};
        if (device_cache) {
            LaunchPlan lr;
            if ((rc = plan_batched(c, MODE_MSE_REF, std::max(c.dm.L1, g.pop), lr)) != SR_OK) return rc;
        }
        rc = device_cache ? dmemo_enqueue(c, g.cur, c.stream, nullptr, kev) : ga_enqueue_eval(c, g.cur, lp, c.stream);
        if (rc) return rc;
        hipError_t he = hipMemcpyAsync(g.last_mse.data(), c.d_mse.p, (size_t)g.pop * 4,
                                       hipMemcpyDeviceToHost, c.stream);
        if (he == hipSuccess) he = hipStreamSynchronize(c.stream);
        if (he != hipSuccess) return c.hip_fail(he, "GA evaluation");
        rc = ga_status(c);
        if (rc) return rc;
        if (st) {
            std::vector<int32_t> l(g.pop);
            hipMemcpy(l.data(), g.len[g.cur].p, (size_t)g.pop * 4, hipMemcpyDeviceToHost);
            *st = sr_stats{};
            for (int v : l) st->logical_node_evals += (int64_t)v * c.N;
            st->executed_node_evals = st->logical_node_evals;
            st->num_kernel_launches = 5;
            st->total_eval_time_ms =
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            if (device_cache) {
                long long after[9];
                if ((rc = dmemo_stats(c, after)) != SR_OK) return rc;
                float t1 = 0.f, t2 = 0.f;
                hipEventElapsedTime(&t1, kev[0], kev[1]);
                hipEventElapsedTime(&t2, kev[2], kev[3]);
                st->interp_kernel_ms = t1 + t2;
                st->interp_launches = 2;
                st->executed_node_evals = (after[1] - before[1]) * c.N;
                st->column_reads = (after[2] - before[2]) * c.N;
                st->column_writes = (after[3] - before[3]) * c.N;
                st->executed_op_evals = (after[4] - before[4]) * c.N;
                st->num_cached = (int32_t)after[7];
                st->num_subtrees = (int32_t)after[7];
            }
        }
    } else {
        std::vector<int32_t> tok, off, len;
        std::vector<double> val;
        rc = ga_download(c, tok, val, off, len);
        if (rc) return rc;
        sr_population pop{g.pop, tok.data(), val.data(), nullptr, off.data(), len.data()};
        rc = sr_eval_mse(x, &pop, mode, g.last_mse.data(), nullptr, st);
        if (rc) return rc;
    }
    if (mse_out) std::memcpy(mse_out, g.last_mse.data(), (size_t)g.pop * 4);
    return SR_OK;
}

int sr_ga_next(sr_ctx *x, const sr_select_params *spp, const float *fitness, uint64_t seed, uint32_t gen)
{
    if (!x || !spp) return SR_ERR_INVALID_ARG;
    Ctx &c = x->c;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    hipSetDevice(c.device);
    int rc = ga_ready(c);
    if (rc) return rc;
    GaState &g = c.ga;
    const float *fit = fitness;
    if (!fit) {
        if ((int)g.last_mse.size() != g.pop) return c.fail(SR_ERR_INVALID_ARG, "no fitness: evaluate first");
        fit = g.last_mse.data();
    }
    SelectParams sp = to_select(*spp, g.pop);
    hipError_t he = hipMemcpyAsync(g.fit.p, fit, (size_t)g.pop * 4, hipMemcpyHostToDevice, c.stream);
    if (he == hipSuccess) he = launch_u32_set(g.gen.as<uint32_t>(), gen, c.stream);
    if (he != hipSuccess) return c.hip_fail(he, "GA next setup");
    rc = ga_enqueue_next(c, g.cur, sp, seed, g.fit.as<float>(), c.stream);
    if (rc) return rc;
    if ((he = hipStreamSynchronize(c.stream)) != hipSuccess) return c.hip_fail(he, "GA next");
    rc = ga_status(c);
    if (rc) return rc;
    g.cur ^= 1;
    g.last_mse.clear();
    return SR_OK;
}

int sr_ga_run(sr_ctx *x, const sr_select_params *spp, int mode, int32_t gens, uint32_t gen0, uint64_t seed,
              int use_graph, float *mse_hist, double *ms_out, sr_stats *st)
{
    if (!x || !spp || gens < 0) return SR_ERR_INVALID_ARG;
    Ctx &c = x->c;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    hipSetDevice(c.device);
    int rc = ga_ready(c);
    if (rc) return rc;
    if (!c.X) return c.fail(SR_ERR_NO_DATA, "no data: call sr_set_data first");
    GaState &g = c.ga;
    sr_stats acc{};
    const SelectParams sp = to_select(*spp, g.pop);
    hipError_t he;
    const bool device_cache = mode == SR_EVAL_CACHED && c.memo.device;
    if (mode != SR_EVAL_BATCHED && !device_cache) {
        // Host-assisted modes: evaluate through the host path, vary on the device.
        auto t0 = std::chrono::steady_clock::now();
        std::vector<float> m(g.pop);
        for (int k = 0; k < gens; ++k) {
            sr_stats s1{};
            rc = sr_ga_evaluate(x, mode, m.data(), &s1);
            if (rc) return rc;
            acc.logical_node_evals += s1.logical_node_evals;
            acc.executed_node_evals += s1.executed_node_evals;
            acc.detect_time_ms += s1.detect_time_ms;
            acc.kernel_time_ms += s1.kernel_time_ms;
            acc.interp_kernel_ms += s1.interp_kernel_ms;
            acc.interp_launches += s1.interp_launches;
            acc.executed_op_evals += s1.executed_op_evals;
            acc.column_reads += s1.column_reads;
            acc.column_writes += s1.column_writes;
            acc.transfer_time_ms += s1.transfer_time_ms;
            acc.jit_compile_time_ms += s1.jit_compile_time_ms;
            acc.num_kernel_launches += s1.num_kernel_launches + 2;
            acc.num_new_subtrees += s1.num_new_subtrees;
            acc.num_cached = s1.num_cached;
            if (mse_hist) std::memcpy(mse_hist + (size_t)k * g.pop, m.data(), (size_t)g.pop * 4);
            rc = sr_ga_next(x, spp, m.data(), seed, gen0 + (uint32_t)k);
            if (rc) return rc;
        }
        double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (ms_out) *ms_out = ms;
        acc.total_eval_time_ms = ms;
        if (st) *st = acc;
        return SR_OK;
    }

    LaunchPlan lp;
    rc = plan_batched(c, MODE_MSE, g.pop, lp);
    if (rc) return rc;
    long long dm_before[9] = {0};
    if (device_cache) {
        if ((rc = dmemo_setup(c, c.memo.capacity, c.memo.min_size)) != SR_OK) return rc;
        if ((rc = dmemo_stats(c, dm_before)) != SR_OK) return rc;
        LaunchPlan lr;  // skeleton launches: buffers (spill area) sized before any capture
        if ((rc = plan_batched(c, MODE_MSE_REF, std::max(c.dm.L1, g.pop), lr)) != SR_OK) return rc;
    }
    if ((he = g.hist.ensure((size_t)std::max(gens, 1) * g.pop * 4)) != hipSuccess) return c.hip_fail(he, "hipMalloc");
    if ((he = g.hist_len.ensure((size_t)std::max(gens, 1) * g.pop * 4)) != hipSuccess)
        return c.hip_fail(he, "hipMalloc");
    // The null stream cannot be captured: a graph run then goes on the
    // context's own stream, after the caller's null-stream work has drained.
    hipStream_t s = c.stream;
    if (use_graph && !s) {
        if ((he = hipStreamSynchronize(nullptr)) != hipSuccess) return c.hip_fail(he, "sync");
        s = c.own_stream;
    }
    // without a graph, the device cache's two interpreter launches are timed
    // with events (one synchronisation per generation)
    const bool time_launches = device_cache && !use_graph;
    hipEvent_t kev[4] = {c.ev[4], c.ev[5], c.ev2[4], c.ev2[5]};
    auto enqueue_gen = [&](int b) -> int {
        int r = device_cache ? dmemo_enqueue(c, b, s, nullptr, time_launches ? kev : nullptr)
                             : ga_enqueue_eval(c, b, lp, s);
        if (r) return r;
        // the ranking launch also writes this generation's history row
        GaRecord rec;
        rec.hist = g.hist.as<uint32_t>();
        rec.hist_len = g.hist_len.as<uint32_t>();
        rec.len = g.len[b].as<int32_t>();
        rec.gen = g.gen.as<uint32_t>();
        rec.gen0 = gen0;
        rec.max_gens = gens;
        r = ga_enqueue_next(c, b, sp, seed, c.d_mse.as<float>(), s, rec);
        if (r) return r;
        hipError_t e = launch_u32_add(g.gen.as<uint32_t>(), 1u, s);
        if (e != hipSuccess) return c.hip_fail(e, "gen counter");
        return SR_OK;
    };

    if (use_graph) {
        uint64_t key = 0;
        // every pointer a captured node holds: a buffer reallocated since the
        // capture (any of them, whichever grew) forces a new capture
        const void *ptrs[] = {g.tok[0].p, g.tok[1].p, g.val[0].p, g.val[1].p, g.len[0].p, g.len[1].p,
                              g.progs.p, g.prog_len.p, g.size_s.p, g.need_s.p, g.frames.p, g.order_len.p,
                              g.order_fit.p, g.status.p, g.max_need.p, g.gen.p, g.hist.p, g.hist_len.p,
                              c.X, c.d_mse.p, c.d_cnt.p, c.d_psum.p, c.d_pcnt.p, c.d_counters.p, (const void *)s,
                              device_cache ? c.dm.cols.p : nullptr, device_cache ? c.dm.ctr.p : nullptr, c.d_spill.p};
        for (const void *q : ptrs) key = mix(key, (uint64_t)(uintptr_t)q);
        for (uint64_t v : {(uint64_t)g.pop, (uint64_t)g.cap, (uint64_t)c.N, (uint64_t)c.nv, (uint64_t)lp.dpt,
                           (uint64_t)lp.chunk, (uint64_t)lp.nq, (uint64_t)lp.grid, (uint64_t)lp.P, seed,
                           (uint64_t)sp.tournament, (uint64_t)sp.elites, (uint64_t)gen0, (uint64_t)gens,
                           (uint64_t)device_cache, (uint64_t)lp.dpt, (uint64_t)c.tune.cache_dpt,
                           (uint64_t)c.tune.kernel, (uint64_t)c.tune.cache_kernel})
            key = mix(key, v);
        uint32_t pm;
        std::memcpy(&pm, &sp.p_mutation, 4);
        key = mix(key, pm);
        if (key != g.graph_key) {
            g.release_graphs();
            for (int b = 0; b < 2; ++b) {
                hipGraph_t graph = nullptr;
                if ((he = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal)) != hipSuccess)
                    return c.hip_fail(he, "begin capture");
                rc = enqueue_gen(b);
                he = hipStreamEndCapture(s, &graph);
                if (rc) {
                    if (graph) hipGraphDestroy(graph);
                    return rc;
                }
                if (he != hipSuccess) return c.hip_fail(he, "end capture");
                he = hipGraphInstantiate(&g.graph[b], graph, nullptr, nullptr, 0);
                hipGraphDestroy(graph);
                if (he != hipSuccess) return c.hip_fail(he, "graph instantiate");
            }
            g.graph_key = key;
            acc.graph_captures += 2;
        }
    }
    if ((he = launch_u32_set(g.gen.as<uint32_t>(), gen0, s)) != hipSuccess) return c.hip_fail(he, "gen set");
    if ((he = hipStreamSynchronize(s)) != hipSuccess) return c.hip_fail(he, "sync");
    hipEventRecord(c.ev[0], s);
    for (int k = 0; k < gens; ++k) {
        if (use_graph) {
            if ((he = hipGraphLaunch(g.graph[g.cur], s)) != hipSuccess) return c.hip_fail(he, "graph launch");
        } else {
            rc = enqueue_gen(g.cur);
            if (rc) return rc;
            if (time_launches) {
                if ((he = hipEventSynchronize(kev[3])) != hipSuccess) return c.hip_fail(he, "device cache");
                float t1 = 0.f, t2 = 0.f;
                hipEventElapsedTime(&t1, kev[0], kev[1]);
                hipEventElapsedTime(&t2, kev[2], kev[3]);
                acc.interp_kernel_ms += t1 + t2;
            }
        }
        g.cur ^= 1;
    }
    hipEventRecord(c.ev[1], s);
    if ((he = hipStreamSynchronize(s)) != hipSuccess) return c.hip_fail(he, "GA run");
    float ms = 0.f;
    hipEventElapsedTime(&ms, c.ev[0], c.ev[1]);
    rc = ga_status(c);
    if (rc) return rc;
    std::vector<int32_t> lens((size_t)gens * g.pop);
    if (gens > 0) {
        if (mse_hist &&
            (he = hipMemcpy(mse_hist, g.hist.p, (size_t)gens * g.pop * 4, hipMemcpyDeviceToHost)) != hipSuccess)
            return c.hip_fail(he, "D2H history");
        if ((he = hipMemcpy(lens.data(), g.hist_len.p, lens.size() * 4, hipMemcpyDeviceToHost)) != hipSuccess)
            return c.hip_fail(he, "D2H lengths");
    }
    for (int v : lens) acc.logical_node_evals += (int64_t)v * c.N;
    acc.executed_node_evals = acc.logical_node_evals;
    if (device_cache) {
        long long after[9];
        if ((rc = dmemo_stats(c, after)) != SR_OK) return rc;
        acc.executed_node_evals = (after[1] - dm_before[1]) * c.N;
        acc.column_reads = (after[2] - dm_before[2]) * c.N;
        acc.column_writes = (after[3] - dm_before[3]) * c.N;
        acc.executed_op_evals = (after[4] - dm_before[4]) * c.N;
        acc.num_cached = (int32_t)after[7];
        acc.num_subtrees = (int32_t)after[7];
        acc.num_new_subtrees = (int32_t)(after[3] - dm_before[3]);
        acc.interp_launches = 2 * gens;
    }
    acc.kernel_time_ms = ms;
    acc.total_eval_time_ms = ms;
    acc.num_kernel_launches = gens * 11;
    if (ms_out) *ms_out = ms;
    if (st) *st = acc;
    g.last_mse.clear();
    return SR_OK;
}

}  // extern "C"

extern "C" int sr_tree_splice(sr_ctx *x, int32_t n, int32_t cap, const int32_t *ltok, const float *lval,
                              const int32_t *llen, int32_t n_left, const int32_t *lsel, const int32_t *lidx,
                              const int32_t *rtok, const float *rval, const int32_t *rlen, int32_t n_right,
                              const int32_t *rsel, const int32_t *ridx, int32_t *otok, float *oval,
                              int32_t *olen)
{
    if (!x || n < 0 || cap <= 0 || n_left <= 0 || n_right <= 0) return SR_ERR_INVALID_ARG;
    if (n == 0) return SR_OK;
    if (!ltok || !lval || !llen || !lsel || !lidx || !rtok || !rval || !rlen || !rsel || !ridx || !otok ||
        !oval || !olen)
        return SR_ERR_INVALID_ARG;
    Ctx &c = x->c;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    hipSetDevice(c.device);
    for (int k = 0; k < n; ++k)
        if (lsel[k] < 0 || lsel[k] >= n_left || rsel[k] < 0 || rsel[k] >= n_right)
            return c.fail(SR_ERR_INVALID_ARG, "parent selector out of range");
    for (int i = 0; i < n_left; ++i)
        if (llen[i] < 0 || llen[i] > cap) return c.fail(SR_ERR_INVALID_ARG, "left length outside 0..cap");
    for (int i = 0; i < n_right; ++i)
        if (rlen[i] < 0 || rlen[i] > cap) return c.fail(SR_ERR_INVALID_ARG, "right length outside 0..cap");
    const size_t L = (size_t)n_left * cap, R = (size_t)n_right * cap, O = (size_t)n * cap;
    DevBuf d;
    const size_t bytes = (L + R + O) * 8 + ((size_t)n_left + n_right + 5 * (size_t)n) * 4;
    hipError_t he = d.ensure(bytes);
    if (he != hipSuccess) return c.hip_fail(he, "hipMalloc splice");
    char *p = d.as<char>();
    int32_t *dlt = (int32_t *)p; p += L * 4;
    float *dlv = (float *)p; p += L * 4;
    int32_t *drt = (int32_t *)p; p += R * 4;
    float *drv = (float *)p; p += R * 4;
    int32_t *dot = (int32_t *)p; p += O * 4;
    float *dov = (float *)p; p += O * 4;
    int32_t *dll = (int32_t *)p; p += (size_t)n_left * 4;
    int32_t *drl = (int32_t *)p; p += (size_t)n_right * 4;
    int32_t *dls = (int32_t *)p; p += (size_t)n * 4;
    int32_t *dli = (int32_t *)p; p += (size_t)n * 4;
    int32_t *drs = (int32_t *)p; p += (size_t)n * 4;
    int32_t *dri = (int32_t *)p; p += (size_t)n * 4;
    int32_t *dol = (int32_t *)p;
    hipStream_t s = c.stream;
    struct Cp { void *dst; const void *src; size_t b; } ups[] = {
        {dlt, ltok, L * 4}, {dlv, lval, L * 4}, {drt, rtok, R * 4}, {drv, rval, R * 4},
        {dll, llen, (size_t)n_left * 4}, {drl, rlen, (size_t)n_right * 4}, {dls, lsel, (size_t)n * 4},
        {dli, lidx, (size_t)n * 4}, {drs, rsel, (size_t)n * 4}, {dri, ridx, (size_t)n * 4}};
    for (auto &u : ups)
        if ((he = hipMemcpyAsync(u.dst, u.src, u.b, hipMemcpyHostToDevice, s)) != hipSuccess) break;
    if (he == hipSuccess)
        he = launch_splice(n, cap, dlt, dlv, dll, dls, dli, drt, drv, drl, drs, dri, dot, dov, dol, s);
    if (he == hipSuccess) he = hipMemcpyAsync(otok, dot, O * 4, hipMemcpyDeviceToHost, s);
    if (he == hipSuccess) he = hipMemcpyAsync(oval, dov, O * 4, hipMemcpyDeviceToHost, s);
    if (he == hipSuccess) he = hipMemcpyAsync(olen, dol, (size_t)n * 4, hipMemcpyDeviceToHost, s);
    if (he == hipSuccess) he = hipStreamSynchronize(s);
    d.release();
    if (he != hipSuccess) return c.hip_fail(he, "splice");
    return SR_OK;
}

// ---- evogp population layout (defs.h:30-107) ---------------------------------
extern "C" int sr_evogp_to_prefix(sr_ctx *x, int32_t pop, int32_t max_len, const float *d_value,
                                  const int16_t *d_type, const int16_t *d_size, int32_t cap, int32_t *d_tok,
                                  float *d_val, int32_t *d_len)
{
    if (!x || pop < 0 || max_len <= 0 || cap <= 0 || (pop > 0 && (!d_value || !d_type || !d_size || !d_tok ||
                                                                  !d_val || !d_len)))
        return SR_ERR_INVALID_ARG;
    Ctx &c = x->c;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    hipSetDevice(c.device);
    if (pop == 0) return SR_OK;
    hipError_t he = launch_evogp_to_prefix(pop, max_len, d_value, d_type, d_size, cap, d_tok, d_val, d_len, c.stream);
    if (he == hipSuccess) he = hipStreamSynchronize(c.stream);
    if (he != hipSuccess) return c.hip_fail(he, "evogp to prefix");
    return SR_OK;
}

extern "C" int sr_prefix_to_evogp(sr_ctx *x, int32_t pop, int32_t cap, const int32_t *d_tok, const float *d_val,
                                  const int32_t *d_len, int32_t max_len, float *d_value, int16_t *d_type,
                                  int16_t *d_size)
{
    if (!x || pop < 0 || max_len <= 0 || cap <= 0 || (pop > 0 && (!d_value || !d_type || !d_size || !d_tok ||
                                                                  !d_val || !d_len)))
        return SR_ERR_INVALID_ARG;
    Ctx &c = x->c;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    hipSetDevice(c.device);
    if (pop == 0) return SR_OK;
    hipError_t he = launch_prefix_to_evogp(pop, cap, d_tok, d_val, d_len, max_len, d_value, d_type, d_size, c.stream);
    if (he == hipSuccess) he = hipStreamSynchronize(c.stream);
    if (he != hipSuccess) return c.hip_fail(he, "prefix to evogp");
    return SR_OK;
}

extern "C" int sr_ga_import_evogp(sr_ctx *x, int32_t pop, int32_t max_len, const float *d_value,
                                  const int16_t *d_type, const int16_t *d_size)
{
    if (!x || pop <= 0 || max_len <= 0 || !d_value || !d_type || !d_size) return SR_ERR_INVALID_ARG;
    Ctx &c = x->c;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    hipSetDevice(c.device);
    int rc = ga_ready(c);
    if (rc) return rc;
    const int cap = c.ga.p.max_tokens + c.ga.p.max_depth + 2;
    rc = ga_alloc(c, pop, cap);
    if (rc) return rc;
    GaState &g = c.ga;
    hipError_t he = launch_evogp_to_prefix(pop, max_len, d_value, d_type, d_size, cap, g.tok[g.cur].as<int32_t>(),
                                           g.val[g.cur].as<float>(), g.len[g.cur].as<int32_t>(), c.stream);
    std::vector<int32_t> len(pop);
    if (he == hipSuccess)
        he = hipMemcpyAsync(len.data(), g.len[g.cur].p, (size_t)pop * 4, hipMemcpyDeviceToHost, c.stream);
    if (he == hipSuccess) he = hipStreamSynchronize(c.stream);
    if (he != hipSuccess) return c.hip_fail(he, "evogp import");
    for (int i = 0; i < pop; ++i)
        if (len[i] < 1)
            return c.fail(SR_ERR_BAD_EXPR, "individual " + std::to_string(i) + ": evogp tree size outside 1..cap");
    return SR_OK;
}

extern "C" int sr_ga_export_evogp(sr_ctx *x, int32_t max_len, float *d_value, int16_t *d_type, int16_t *d_size)
{
    if (!x || max_len <= 0 || !d_value || !d_type || !d_size) return SR_ERR_INVALID_ARG;
    Ctx &c = x->c;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    hipSetDevice(c.device);
    int rc = ga_ready(c);
    if (rc) return rc;
    GaState &g = c.ga;
    hipError_t he = launch_prefix_to_evogp(g.pop, g.cap, g.tok[g.cur].as<int32_t>(), g.val[g.cur].as<float>(),
                                           g.len[g.cur].as<int32_t>(), max_len, d_value, d_type, d_size, c.stream);
    if (he == hipSuccess) he = hipStreamSynchronize(c.stream);
    if (he != hipSuccess) return c.hip_fail(he, "evogp export");
    return SR_OK;
}

// One select + vary step on host arrays (the evolve step of the drop-in
// boundary): the packed population and its fitness in, the next generation
// out in stride-cap layout.  Uses the context's GA storage.
extern "C" int sr_evolve_step(sr_ctx *x, const sr_population *pop_in, const float *fitness, const sr_ga_params *p,
                              const sr_select_params *sp, uint64_t seed, uint32_t gen, int32_t cap_out,
                              int32_t *tok_out, float *val_out, int32_t *len_out)
{
    if (!x || !pop_in || !fitness || !p || !sp || !tok_out || !val_out || !len_out || pop_in->num_exprs <= 0 ||
        !pop_in->tokens || !pop_in->offsets || !pop_in->lengths || (!pop_in->values && !pop_in->values_f32))
        return SR_ERR_INVALID_ARG;
    Ctx &c = x->c;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    hipSetDevice(c.device);
    int rc = ga_configure(c, p);
    if (rc) return rc;
    const int cap = p->max_tokens + p->max_depth + 2;
    if (cap_out != cap) return c.fail(SR_ERR_INVALID_ARG, "cap_out must equal max_tokens + max_depth + 2");
    rc = validate_population(c, *pop_in);
    if (rc) return rc;
    const int E = pop_in->num_exprs;
    std::vector<int32_t> tok((size_t)E * cap, 0), len(E);
    std::vector<float> val((size_t)E * cap, 0.0f);
    for (int e = 0; e < E; ++e) {
        const int32_t L = pop_in->lengths[e], o = pop_in->offsets[e];
        if (L > cap) return c.fail(SR_ERR_INVALID_ARG, "expression " + std::to_string(e) + " longer than cap");
        len[e] = L;
        for (int k = 0; k < L; ++k) {
            tok[(size_t)e * cap + k] = pop_in->tokens[o + k];
            val[(size_t)e * cap + k] = pop_in->values ? (float)pop_in->values[o + k] : pop_in->values_f32[o + k];
        }
    }
    rc = ga_alloc(c, E, cap);
    if (rc) return rc;
    c.ga.cur = 0;
    c.ga.ready = true;
    if ((rc = sr_ga_set_population(x, tok.data(), val.data(), len.data(), E, cap))) return rc;
    if ((rc = sr_ga_next(x, sp, fitness, seed, gen))) return rc;
    int32_t np = 0, nc = 0;
    return sr_ga_get_population(x, tok_out, val_out, len_out, &np, &nc);
}
