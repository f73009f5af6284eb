// sr_host.cpp -- context, data upload, program build and the plain / batched
// evaluation paths behind the C ABI (include/sr_hip.h).
//
// Evaluation flow (the reference's evaluate_gpu_simple, cl_gpu_simple.cu:187-310,
// without its materialised sq_err[E*N] buffer and separate reduce launch):
//   host : validate + compile each expression        (O(total tokens))
//   H2D  : one async copy of the packed program from pinned staging
//   GPU  : interpreter with fused squared-error partials -> k_finalize -> mse[E]
//   D2H  : mse[E] (+ valid counts)
#include <algorithm>
#include <chrono>
#include <cmath>
#include <numeric>

#include "sr_adjoint_prog.h"
#include "sr_compile_wave.h"
#include "sr_ctx.h"
#include "sr_kernels.h"
#include "sr_launch.h"

namespace sr {

void expr_values_f32(const sr_population &pop, int e, std::vector<float> &out)
{
    const int32_t o = pop.offsets[e], L = pop.lengths[e];
    out.resize(L);
    if (pop.values) {
        for (int k = 0; k < L; ++k) out[k] = (float)pop.values[o + k];
    } else {
        for (int k = 0; k < L; ++k) out[k] = pop.values_f32[o + k];
    }
}

int build_program(Ctx &c, const sr_population &pop, bool raw, Program &out, int num_slots)
{
    const int E = pop.num_exprs;
    out.ins.clear();
    out.off.assign(E, 0);
    out.len.assign(E, 0);
    out.need.assign(E, 0);
    out.max_need = 0;
    out.tokens = 0;
    out.ops = out.refs = out.stores = 0;
    out.ref_pf.assign(num_slots > 0 ? (size_t)E * kRefPrefetch : 0, -1);
    std::vector<float> vals;
    std::vector<int> size(MAX_EXPR_LEN), need(MAX_EXPR_LEN);
    std::vector<CompileFrame> frames(MAX_EXPR_LEN + 1);
    std::vector<Ins> tmp(MAX_EXPR_LEN);
    for (int e = 0; e < E; ++e) {
        const int32_t L = pop.lengths[e];
        const int32_t *tok = pop.tokens + pop.offsets[e];
        if (L <= 0 || L > MAX_EXPR_LEN)
            return c.fail(SR_ERR_BAD_EXPR, "expression " + std::to_string(e) + ": length " +
                                               std::to_string(L) + " outside 1.." +
                                               std::to_string(MAX_EXPR_LEN));
        int depth = raw_stack_depth(tok, L, num_slots > 0);
        if (depth < 0)
            return c.fail(SR_ERR_BAD_EXPR,
                          "expression " + std::to_string(e) + ": malformed prefix token stream");
        expr_values_f32(pop, e, vals);
        for (int k = 0; k < L; ++k) {
            if (tok[k] == TOK_VAR && !(vals[k] >= 0.0f))
                return c.fail(SR_ERR_BAD_EXPR, "expression " + std::to_string(e) +
                                                   ": negative variable index");
            // OP_REF only appears in skeletons built by the cache, whose slots
            // all hold a value column; anywhere else it is an invalid token.
            if ((tok[k] == OP_REF || tok[k] == OP_STORE) && !(vals[k] >= 0.0f && vals[k] < (float)num_slots))
                return c.fail(SR_ERR_BAD_EXPR, "expression " + std::to_string(e) +
                                                   ": cache reference outside the cache");
        }
        out.off[e] = (int32_t)out.ins.size();
        out.tokens += L;
        for (int k = 0; k < L; ++k) {
            out.ops += tok[k] > 0;
            out.refs += tok[k] == OP_REF;
            out.stores += tok[k] == OP_STORE;
        }
        if (raw) {
            for (int k = 0; k < L; ++k) {
                Ins ins;
                lower_token(tok[k], vals[k], c.nv, &ins);
                out.ins.push_back(ins);
            }
            out.len[e] = L;
            out.need[e] = depth;
        } else {
            int sn = 0;
            int n = compile_expr(tok, vals.data(), L, c.nv, size.data(), need.data(), frames.data(),
                                 tmp.data(), &sn);
            if (n < 0)
                return c.fail(SR_ERR_BAD_EXPR, "expression " + std::to_string(e) + ": compile failed");
            if (num_slots > 0) {
                // the first kRefPrefetch OP_REF operands are loaded a tile ahead
                int j = 0;
                for (int k = 0; k < n && j < kRefPrefetch; ++k) {
                    if (bc_src(tmp[k].code) != SRC_REF) continue;
                    tmp[k].code |= (j + 1) << BC_PF_SHIFT;
                    out.ref_pf[(size_t)e * kRefPrefetch + j] = tmp[k].arg;
                    ++j;
                }
            }
            out.ins.insert(out.ins.end(), tmp.begin(), tmp.begin() + n);
            out.len[e] = n;
            out.need[e] = sn;
        }
        out.max_need = std::max(out.max_need, out.need[e]);
    }
    // Longest first: the persistent queue then approximates LPT scheduling.
    out.order.resize(E);
    std::iota(out.order.begin(), out.order.end(), 0);
    std::stable_sort(out.order.begin(), out.order.end(),
                     [&](int a, int b) { return out.len[a] > out.len[b]; });
    return SR_OK;
}

static int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

int plan_launch(Ctx &c, int mode, LaunchPlan &lp, bool allow_vec)
{
    lp.kind = KERNEL_LANE;
    // (the vector interpreter addresses columns with 32-bit element offsets)
    if (allow_vec && c.N <= (int64_t)UINT32_MAX &&
        ((mode == MODE_MSE && c.tune.kernel == KERNEL_VEC) ||
         (mode == MODE_MSE_REF && c.tune.cache_kernel == KERNEL_VEC)))
        lp.kind = KERNEL_VEC;
    lp.xc = pick_xc(c.nv);
    lp.dpt = mode == MODE_MSE_REF ? (c.tune.cache_dpt > 0 ? c.tune.cache_dpt : 3) : c.tune.dpt;
    if (lp.dpt < 1 || lp.dpt > 4) lp.dpt = 4;
    if (lp.dpt == 3 && lp.xc != 4) lp.dpt = 4;  // dpt 3 is built for <= 4 variables only
    lp.xc = pick_xc_batched(c.nv, lp.dpt);
    if (lp.xc == 32 && lp.dpt > 2) lp.dpt = 2;
    lp.nq = c.tune.n_queues > 1 ? c.tune.n_queues : 1;
    // Chunks are multiples of 768 = lcm(64 * dpt) and queue slices of 1536,
    // so the datapoints of every (expression, chunk) partial sum, and the
    // order each lane adds them in (lane + 64 k, k ascending), do not depend
    // on dpt: MSEs are bit-identical across dpt for equal chunk and queues.
    constexpr int64_t kChunkQuantum = 768;
    lp.slice = round_up((c.N + lp.nq - 1) / lp.nq, 2 * kChunkQuantum);
    lp.chunk = round_up(std::max<int64_t>(c.tune.chunk, kChunkQuantum), kChunkQuantum);
    if (lp.chunk > lp.slice) lp.chunk = lp.slice;
    lp.cpq = (int)((lp.slice + lp.chunk - 1) / lp.chunk);
    lp.P = lp.nq * lp.cpq;
    int occ = lp.kind == KERNEL_VEC ? vec_occupancy(mode) : batched_occupancy(mode, lp.dpt, lp.xc);
    if (occ <= 0) return c.fail(SR_ERR_UNSUPPORTED, "no kernel variant for this dpt / variable count");
    int bpc = c.tune.blocks_per_cu;
    if (bpc <= 0 || bpc > occ) bpc = occ;
    lp.grid = (int)round_up((int64_t)c.num_cu * bpc, lp.nq);
    return SR_OK;
}

int plan_batched(Ctx &c, int mode, int E, LaunchPlan &lp)
{
    hipError_t he;
    int rc = plan_launch(c, mode, lp);
    if (rc) return rc;
    // queue counters: two sets of 64 (the device cache's two launches use one each)
    if ((he = c.d_counters.ensure(2 * 64 * 4)) != hipSuccess) return c.hip_fail(he, "hipMalloc");
    if (lp.kind == KERNEL_VEC &&
        (he = c.d_spill.ensure((size_t)lp.grid * (kBlock / kWave) * kVecSpillBytesPerWave)) != hipSuccess)
        return c.hip_fail(he, "hipMalloc");
    if (is_mse_mode(mode)) {
        if ((he = c.d_psum.ensure((size_t)E * lp.P * 8)) != hipSuccess) return c.hip_fail(he, "hipMalloc");
        if ((he = c.d_pcnt.ensure((size_t)E * lp.P * 8)) != hipSuccess) return c.hip_fail(he, "hipMalloc");
        if ((he = c.d_mse.ensure((size_t)E * 4)) != hipSuccess) return c.hip_fail(he, "hipMalloc");
        if ((he = c.d_cnt.ensure((size_t)E * 8)) != hipSuccess) return c.hip_fail(he, "hipMalloc");
    }
    return SR_OK;
}

int enqueue_batched(Ctx &c, int mode, const LaunchPlan &lp, BatchArgs a, int E, hipStream_t s,
                    hipEvent_t k0, hipEvent_t k1)
{
    EvalBufs b{c.d_psum.as<double>(), c.d_pcnt.as<long long>(), c.d_counters.as<uint32_t>(),
               c.d_mse.as<float>(), c.d_cnt.as<long long>()};
    b.spill = c.d_spill.as<float>();
    return enqueue_batched(c, mode, lp, a, E, s, b, k0, k1);
}

int enqueue_batched(Ctx &c, int mode, const LaunchPlan &lp, BatchArgs a, int E, hipStream_t s,
                    const EvalBufs &b, hipEvent_t k0, hipEvent_t k1)
{
    hipError_t he;
    a.X = c.X;
    a.psum = b.psum;
    a.pcnt = b.pcnt;
    a.counters = b.counters;
    a.N = c.N;
    a.slice = lp.slice;
    a.chunk = lp.chunk;
    a.n_exprs = E;
    a.nv = c.nv;
    a.n_queues = lp.nq;
    a.chunks_per_queue = lp.cpq;
    a.P = lp.P;
    if (!b.counters_zeroed && (he = launch_zero_u32(a.counters, lp.nq, s)) != hipSuccess)
        return c.hip_fail(he, "zero counters");
    if (k0) hipEventRecord(k0, s);
    a.spill = b.spill;
    // no more blocks than the work items can keep busy (a wave takes one
    // item; small launches would otherwise start thousands of idle waves)
    const int64_t per_queue = ((int64_t)E * lp.cpq + kBlock / kWave - 1) / (kBlock / kWave);
    const int grid = (int)std::min<int64_t>(lp.grid, (int64_t)lp.nq * std::max<int64_t>(per_queue, 1));
    he = lp.kind == KERNEL_VEC ? launch_vec(mode, a, grid, s) : launch_batched(mode, lp.dpt, lp.xc, a, grid, s);
    if (he != hipSuccess) return c.hip_fail(he, "batched launch");
    if (k1) hipEventRecord(k1, s);
    if (is_mse_mode(mode) && b.finalize) {
        if ((he = launch_finalize(a.psum, a.pcnt, lp.P, E, b.mse, b.cnt, s, b.fin_count)) != hipSuccess)
            return c.hip_fail(he, "finalize launch");
    }
    return SR_OK;
}

int run_batched(Ctx &c, const Program &p, int mode, const float *cache_cols, float *out_dev,
                const int32_t *out_col_host, sr_stats *st, int set, bool wait)
{
    HostBuf &h_prog = set ? c.h_prog2 : c.h_prog;
    HostBuf &h_meta = set ? c.h_meta2 : c.h_meta;
    DevBuf &d_prog = set ? c.d_prog2 : c.d_prog;
    DevBuf &d_meta = set ? c.d_meta2 : c.d_meta;
    hipEvent_t *ev = set ? c.ev2 : c.ev;
    const int E = (int)p.len.size();
    if (E == 0) return SR_OK;
    if (p.max_need > kStackSlots)
        return c.fail(SR_ERR_UNSUPPORTED, "expression needs more than 8 stack slots");
    hipError_t he;
    LaunchPlan lp;
    int rc = plan_batched(c, mode, E, lp);
    if (rc) return rc;

    // Pack program + metadata into pinned staging, one async copy each.
    const size_t prog_bytes = p.ins.size() * sizeof(Ins);
    // a program built without cache slots has no OP_REF: nothing to prefetch
    const bool pf = has_ref(mode);
    const size_t meta_ints = (size_t)E * (pf ? 4 + kRefPrefetch : 4);
    if ((he = h_prog.ensure(prog_bytes + 16)) != hipSuccess) return c.hip_fail(he, "hipHostMalloc");
    if ((he = h_meta.ensure(meta_ints * 4)) != hipSuccess) return c.hip_fail(he, "hipHostMalloc");
    std::memcpy(h_prog.p, p.ins.data(), prog_bytes);
    int32_t *meta = h_meta.as<int32_t>();
    std::memcpy(meta, p.off.data(), E * 4);
    std::memcpy(meta + E, p.len.data(), E * 4);
    std::memcpy(meta + 2 * E, p.order.data(), E * 4);
    if (out_col_host) std::memcpy(meta + 3 * E, out_col_host, E * 4);
    else std::memset(meta + 3 * E, 0, E * 4);
    if (pf) {
        if (p.ref_pf.size() == (size_t)E * kRefPrefetch)
            std::memcpy(meta + 4 * E, p.ref_pf.data(), (size_t)E * kRefPrefetch * 4);
        else
            std::fill(meta + 4 * E, meta + 4 * E + (size_t)E * kRefPrefetch, -1);
    }
    if ((he = d_prog.ensure(prog_bytes + 16)) != hipSuccess) return c.hip_fail(he, "hipMalloc");
    if ((he = d_meta.ensure(meta_ints * 4)) != hipSuccess) return c.hip_fail(he, "hipMalloc");
    hipStream_t s = c.stream;
    hipEventRecord(ev[0], s);
    if ((he = hipMemcpyAsync(d_prog.p, h_prog.p, prog_bytes, hipMemcpyHostToDevice, s)) != hipSuccess)
        return c.hip_fail(he, "H2D program");
    if ((he = hipMemcpyAsync(d_meta.p, h_meta.p, meta_ints * 4, hipMemcpyHostToDevice, s)) != hipSuccess)
        return c.hip_fail(he, "H2D metadata");
    hipEventRecord(ev[1], s);

    BatchArgs a{};
    a.progs = d_prog.as<Ins>();
    a.prog_off = d_meta.as<int32_t>();
    a.prog_len = a.prog_off + E;
    a.order = a.prog_off + 2 * E;
    a.out_col = a.prog_off + 3 * E;
    a.ref_pf = pf ? a.prog_off + 4 * E : nullptr;
    a.cache = cache_cols;
    a.out = out_dev;
    rc = enqueue_batched(c, mode, lp, a, E, s, st ? ev[4] : nullptr, st ? ev[5] : nullptr);
    if (rc) return rc;
    hipEventRecord(ev[2], s);
    if (st) {
        st->num_kernel_launches += is_mse_mode(mode) ? 3 : 2;
        st->executed_node_evals += p.tokens * c.N;
        st->executed_op_evals += p.ops * c.N;
        st->column_reads += p.refs * c.N;
        st->column_writes += (p.stores + (is_column_mode(mode) ? E : 0)) * c.N;
        if (wait) {
            // every caller synchronises right after; waiting here costs nothing
            if ((he = hipEventSynchronize(ev[5])) != hipSuccess) return c.hip_fail(he, "interpreter");
            add_batched_timing(c, set, st);
        }
    }
    return SR_OK;
}

void add_batched_timing(Ctx &c, int set, sr_stats *st)
{
    hipEvent_t *ev = set ? c.ev2 : c.ev;
    float ms = 0.f;
    hipEventElapsedTime(&ms, ev[4], ev[5]);
    st->interp_kernel_ms += ms;
    st->interp_launches += 1;
}

int validate_population(Ctx &c, const sr_population &pop)
{
    for (int e = 0; e < pop.num_exprs; ++e) {
        const int32_t L = pop.lengths[e];
        const int64_t o = pop.offsets[e];
        const int32_t *tok = pop.tokens + o;
        if (o < 0 || L <= 0 || L > MAX_EXPR_LEN)
            return c.fail(SR_ERR_BAD_EXPR, "expression " + std::to_string(e) + ": length " +
                                               std::to_string(L) + " outside 1.." +
                                               std::to_string(MAX_EXPR_LEN));
        if (raw_stack_depth(tok, L) < 0)
            return c.fail(SR_ERR_BAD_EXPR,
                          "expression " + std::to_string(e) + ": malformed prefix token stream");
        for (int k = 0; k < L; ++k) {
            if (tok[k] == OP_REF)
                return c.fail(SR_ERR_BAD_EXPR, "expression " + std::to_string(e) +
                                                   ": cache reference outside the cache");
            if (tok[k] == TOK_VAR) {
                const float v = pop.values ? (float)pop.values[o + k] : pop.values_f32[o + k];
                if (!(v >= 0.0f))
                    return c.fail(SR_ERR_BAD_EXPR, "expression " + std::to_string(e) +
                                                       ": negative variable index");
            }
        }
    }
    return SR_OK;
}

static int run_plain(Ctx &c, const Program &p, sr_stats *st)
{
    const int E = (int)p.len.size();
    hipError_t he;
    const int xc = pick_xc(c.nv);
    const size_t prog_bytes = p.ins.size() * sizeof(Ins);
    if (p.max_need > 64) return c.fail(SR_ERR_UNSUPPORTED, "plain interpreter: stack depth above 64");
    if ((he = c.h_prog.ensure(prog_bytes + 16)) != hipSuccess) return c.hip_fail(he, "hipHostMalloc");
    std::memcpy(c.h_prog.p, p.ins.data(), prog_bytes);
    const int P = (int)((c.N + kBlock - 1) / kBlock);
    if ((he = c.d_prog.ensure(prog_bytes + 16)) != hipSuccess) return c.hip_fail(he, "hipMalloc");
    if ((he = c.d_psum.ensure((size_t)E * P * 8)) != hipSuccess) return c.hip_fail(he, "hipMalloc");
    if ((he = c.d_pcnt.ensure((size_t)E * P * 8)) != hipSuccess) return c.hip_fail(he, "hipMalloc");
    if ((he = c.d_mse.ensure((size_t)E * 4)) != hipSuccess) return c.hip_fail(he, "hipMalloc");
    if ((he = c.d_cnt.ensure((size_t)E * 8)) != hipSuccess) return c.hip_fail(he, "hipMalloc");
    hipStream_t s = c.stream;
    hipEventRecord(c.ev[0], s);
    if ((he = hipMemcpyAsync(c.d_prog.p, c.h_prog.p, prog_bytes, hipMemcpyHostToDevice, s)) != hipSuccess)
        return c.hip_fail(he, "H2D program");
    hipEventRecord(c.ev[1], s);
    const Ins *dprog = c.d_prog.as<Ins>();
    double *psum = c.d_psum.as<double>();
    long long *pcnt = c.d_pcnt.as<long long>();
    hipEventRecord(c.ev[4], s);
    for (int e = 0; e < E; ++e) {
        int slots = p.need[e] <= 16 ? 16 : 64;
        he = launch_plain(slots, xc, dprog + p.off[e], p.len[e], c.X, c.N, c.nv, psum + (size_t)e * P,
                          pcnt + (size_t)e * P, s);
        if (he != hipSuccess) return c.hip_fail(he, "plain launch");
    }
    hipEventRecord(c.ev[5], s);
    if ((he = launch_finalize(psum, pcnt, P, E, c.d_mse.as<float>(), c.d_cnt.as<long long>(), s)) != hipSuccess)
        return c.hip_fail(he, "finalize launch");
    hipEventRecord(c.ev[2], s);
    if (st) {
        st->num_kernel_launches += E + 1;
        st->executed_node_evals += p.tokens * c.N;
        st->executed_op_evals += p.ops * c.N;
        if ((he = hipEventSynchronize(c.ev[5])) != hipSuccess) return c.hip_fail(he, "plain interpreter");
        float ms = 0.f;
        hipEventElapsedTime(&ms, c.ev[4], c.ev[5]);
        st->interp_kernel_ms += ms;
        st->interp_launches += E;
    }
    return SR_OK;
}

int eval_cached(Ctx &c, const sr_population &pop, sr_stats *st);  // sr_cache.cpp
int eval_memo(Ctx &c, const sr_population &pop, sr_stats *st, float *mse_out, int64_t *valid_out);
int eval_jit(Ctx &c, const sr_population &pop, sr_stats *st);  // sr_jit.cpp

}  // namespace sr

using namespace sr;

static double now_ms()
{
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

extern "C" {

int sr_ctx_create(int device, sr_ctx **out)
{
    if (!out) return SR_ERR_INVALID_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return SR_ERR_NO_DEVICE;
    if (device < 0 || device >= n) return SR_ERR_INVALID_ARG;
    if (hipSetDevice(device) != hipSuccess) return SR_ERR_HIP;
    sr_ctx *x = new sr_ctx();
    Ctx &c = x->c;
    c.device = device;
    hipDeviceGetAttribute(&c.num_cu, hipDeviceAttributeMultiprocessorCount, device);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
        size_t k = 0;
        while (k + 1 < sizeof c.arch && prop.gcnArchName[k] && prop.gcnArchName[k] != ':') {
            c.arch[k] = prop.gcnArchName[k];
            ++k;
        }
        c.arch[k] = '\0';
    }
    if (hipStreamCreateWithFlags(&c.own_stream, hipStreamNonBlocking) != hipSuccess) {
        delete x;
        return SR_ERR_HIP;
    }
    c.stream = c.own_stream;
    for (auto &e : c.ev) hipEventCreate(&e);
    for (auto &e : c.ev2) hipEventCreate(&e);
    if (noalloc_scratch_init(device, &c.noalloc_scratch) != hipSuccess) {
        sr_ctx_destroy(x);
        return SR_ERR_HIP;
    }
    *out = x;
    return SR_OK;
}

int sr_ctx_destroy(sr_ctx *x)
{
    if (!x) return SR_OK;
    Ctx &c = x->c;
    hipSetDevice(c.device);
    hipStreamSynchronize(c.stream);
    for (auto &e : c.ev)
        if (e) hipEventDestroy(e);
    for (auto &e : c.ev2)
        if (e) hipEventDestroy(e);
    if (c.own_stream) hipStreamDestroy(c.own_stream);
    delete x;
    return SR_OK;
}

const char *sr_last_error(const sr_ctx *x) { return x ? x->c.err.c_str() : "null context"; }

int sr_runtime_info(int32_t *built, int32_t *runtime, int32_t *driver)
{
    int rt = 0, drv = 0;
    if (hipRuntimeGetVersion(&rt) != hipSuccess) rt = 0;
    if (hipDriverGetVersion(&drv) != hipSuccess) drv = 0;
    if (built) *built = HIP_VERSION;
    if (runtime) *runtime = rt;
    if (driver) *driver = drv;
    // HIP_VERSION = major * 10^7 + minor * 10^5 + patch
    return rt / 10000000 == HIP_VERSION / 10000000 ? SR_OK : SR_ERR_UNSUPPORTED;
}

int sr_set_stream(sr_ctx *x, void *stream)
{
    if (!x) return SR_ERR_INVALID_ARG;
    std::lock_guard<std::recursive_mutex> g(x->c.mu);
    x->c.stream = (hipStream_t)stream;
    return SR_OK;
}

int sr_use_own_stream(sr_ctx *x)
{
    if (!x) return SR_ERR_INVALID_ARG;
    std::lock_guard<std::recursive_mutex> g(x->c.mu);
    x->c.stream = x->c.own_stream;
    return SR_OK;
}

static int check_shape(Ctx &c, int32_t nv, int64_t N)
{
    if (nv < 0 || nv > MAX_VARS) return c.fail(SR_ERR_INVALID_ARG, "nv must be in 0..32");
    if (N <= 0) return c.fail(SR_ERR_INVALID_ARG, "N must be positive");
    return SR_OK;
}

int sr_set_data(sr_ctx *x, const float *X, int32_t nv, int64_t N)
{
    if (!x || !X) return SR_ERR_INVALID_ARG;
    Ctx &c = x->c;
    std::lock_guard<std::recursive_mutex> g(c.mu);
    hipSetDevice(c.device);
    int rc = check_shape(c, nv, N);
    if (rc) return rc;
    size_t bytes = (size_t)(nv + 1) * (size_t)N * sizeof(float);
    hipError_t he = c.X_own.ensure(bytes);
    if (he != hipSuccess) return c.hip_fail(he, "hipMalloc X");
    he = hipMemcpyAsync(c.X_own.p, X, bytes, hipMemcpyHostToDevice, c.stream);
    if (he == hipSuccess) he = hipStreamSynchronize(c.stream);
    if (he != hipSuccess) return c.hip_fail(he, "H2D X");
    c.X = c.X_own.as<float>();
    c.nv = nv;
    c.N = N;
    c.data_version++;
    return SR_OK;
}

int sr_set_data_device(sr_ctx *x, const float *dX, int32_t nv, int64_t N)
{
    if (!x || !dX) return SR_ERR_INVALID_ARG;
    Ctx &c = x->c;
    std::lock_guard<std::recursive_mutex> g(c.mu);
    int rc = check_shape(c, nv, N);
    if (rc) return rc;
    c.X = dX;
    c.nv = nv;
    c.N = N;
    c.data_version++;
    return SR_OK;
}

static int check_pop(Ctx &c, const sr_population *pop)
{
    if (!pop || pop->num_exprs < 0) return c.fail(SR_ERR_INVALID_ARG, "null population");
    if (pop->num_exprs > 0 &&
        (!pop->tokens || !pop->offsets || !pop->lengths || (!pop->values && !pop->values_f32)))
        return c.fail(SR_ERR_INVALID_ARG, "population arrays missing");
    if (!c.X) return c.fail(SR_ERR_NO_DATA, "no data: call sr_set_data first");
    return SR_OK;
}

int sr_eval_mse(sr_ctx *x, const sr_population *pop, int mode, float *mse_out, int64_t *valid_out,
                sr_stats *st)
{
    if (!x || !mse_out) return SR_ERR_INVALID_ARG;
    Ctx &c = x->c;
    std::lock_guard<std::recursive_mutex> g(c.mu);
    hipSetDevice(c.device);
    int rc = check_pop(c, pop);
    if (rc) return rc;
    sr_stats local{};
    sr_stats *S = st ? st : &local;
    *S = sr_stats{};
    const double t0 = now_ms();
    const int E = pop->num_exprs;
    int64_t logical = 0;
    for (int e = 0; e < E; ++e) logical += (int64_t)pop->lengths[e] * c.N;
    S->logical_node_evals = logical;
    if (E == 0) return SR_OK;

    if (mode == SR_EVAL_JIT) {
        rc = eval_jit(c, *pop, S);
    } else if (mode == SR_EVAL_CACHED && c.memo.device) {
        return c.fail(SR_ERR_UNSUPPORTED, "SR_CACHE_DEVICE evaluates the context's GA population "
                                          "(sr_ga_run / sr_ga_evaluate), not a host population");
    } else if (mode == SR_EVAL_CACHED && c.memo.enabled) {
        // results come back through the cache (some are remembered MSEs)
        rc = eval_memo(c, *pop, S, mse_out, valid_out);
        S->total_eval_time_ms = now_ms() - t0;
        return rc;
    } else if (mode == SR_EVAL_CACHED) {
        rc = eval_cached(c, *pop, S);
    } else if (mode == SR_EVAL_PLAIN || mode == SR_EVAL_BATCHED) {
        Program p;
        rc = build_program(c, *pop, mode == SR_EVAL_PLAIN, p);
        if (rc) return rc;
        rc = mode == SR_EVAL_PLAIN ? run_plain(c, p, S) : run_batched(c, p, MODE_MSE, nullptr, nullptr, nullptr, S);
    } else {
        return c.fail(SR_ERR_INVALID_ARG, "unknown evaluation mode");
    }
    if (rc) return rc;
    hipError_t he;
    hipStream_t s = c.stream;
    if ((he = c.h_out.ensure(((size_t)E * 4 + 7) / 8 * 8 + (size_t)E * 8)) != hipSuccess)
        return c.hip_fail(he, "hipHostMalloc");
    float *hm = c.h_out.as<float>();
    long long *hc = (long long *)(c.h_out.as<char>() + (((size_t)E * 4 + 7) / 8) * 8);
    if ((he = hipMemcpyAsync(hm, c.d_mse.p, (size_t)E * 4, hipMemcpyDeviceToHost, s)) != hipSuccess)
        return c.hip_fail(he, "D2H mse");
    if ((he = hipMemcpyAsync(hc, c.d_cnt.p, (size_t)E * 8, hipMemcpyDeviceToHost, s)) != hipSuccess)
        return c.hip_fail(he, "D2H counts");
    hipEventRecord(c.ev[3], s);
    if ((he = hipStreamSynchronize(s)) != hipSuccess) return c.hip_fail(he, "evaluation");
    std::memcpy(mse_out, hm, (size_t)E * 4);
    if (valid_out)
        for (int e = 0; e < E; ++e) valid_out[e] = hc[e];
    float ms = 0.f;
    hipEventElapsedTime(&ms, c.ev[0], c.ev[1]);
    S->transfer_time_ms += ms;
    hipEventElapsedTime(&ms, c.ev[1], c.ev[2]);
    S->kernel_time_ms += ms;
    hipEventElapsedTime(&ms, c.ev[2], c.ev[3]);
    S->transfer_time_ms += ms;
    S->total_eval_time_ms = now_ms() - t0;
    return SR_OK;
}

int sr_eval_predictions(sr_ctx *x, const sr_population *pop, float *pred_out)
{
    if (!x || !pred_out) return SR_ERR_INVALID_ARG;
    Ctx &c = x->c;
    std::lock_guard<std::recursive_mutex> g(c.mu);
    hipSetDevice(c.device);
    int rc = check_pop(c, pop);
    if (rc) return rc;
    const int E = pop->num_exprs;
    if (E == 0) return SR_OK;
    Program p;
    rc = build_program(c, *pop, false, p);
    if (rc) return rc;
    size_t bytes = (size_t)E * (size_t)c.N * sizeof(float);
    hipError_t he = c.d_pred.ensure(bytes);
    if (he != hipSuccess) return c.hip_fail(he, "hipMalloc predictions");
    rc = run_batched(c, p, MODE_PRED, nullptr, c.d_pred.as<float>(), nullptr, nullptr);
    if (rc) return rc;
    he = hipMemcpyAsync(pred_out, c.d_pred.p, bytes, hipMemcpyDeviceToHost, c.stream);
    if (he == hipSuccess) he = hipStreamSynchronize(c.stream);
    if (he != hipSuccess) return c.hip_fail(he, "D2H predictions");
    return SR_OK;
}

int sr_compile_expr(const int32_t *tokens, const float *values, int32_t len, int32_t nv,
                    int32_t *out_code, int32_t *out_arg, int32_t *stack_need)
{
    if (!tokens || !values || !out_code || !out_arg || !stack_need) return SR_ERR_INVALID_ARG;
    if (len <= 0 || len > MAX_EXPR_LEN || raw_stack_depth(tokens, len) < 0) return SR_ERR_BAD_EXPR;
    std::vector<int> size(len), need(len);
    std::vector<CompileFrame> frames(len + 1);
    std::vector<Ins> ins(len);
    int sn = 0;
    int n = compile_expr(tokens, values, len, nv, size.data(), need.data(), frames.data(), ins.data(),
                         &sn);
    if (n < 0) return SR_ERR_BAD_EXPR;
    for (int i = 0; i < n; ++i) {
        out_code[i] = ins[i].code;
        out_arg[i] = ins[i].arg;
    }
    *stack_need = sn;
    return n;
}

int sr_compile_skeleton(const int32_t *tokens, const float *values, int32_t len, int32_t nv, int32_t algorithm,
                        int32_t *out_code, int32_t *out_arg, int32_t *stack_need)
{
    if (!tokens || !values || !out_code || !out_arg || !stack_need || (algorithm != 0 && algorithm != 1))
        return SR_ERR_INVALID_ARG;
    if (len <= 0 || len > MAX_EXPR_LEN) return SR_ERR_BAD_EXPR;
    std::vector<Ins> ins(len);
    int sn = 0, n = -1;
    int32_t pf[kRefPrefetch];
    if (algorithm == 0) {
        if (raw_stack_depth(tokens, len, true) < 0) return SR_ERR_BAD_EXPR;
        std::vector<int> size(len), need(len);
        std::vector<CompileFrame> frames(len + 1);
        n = compile_expr(tokens, values, len, nv, size.data(), need.data(), frames.data(), ins.data(), &sn);
        // the first kRefPrefetch OP_REF operands carry their order + 1
        for (int i = 0, j = 0; n > 0 && i < n && j < kRefPrefetch; ++i)
            if (bc_src(ins[i].code) == SRC_REF) ins[i].code |= ++j << BC_PF_SHIFT;
    } else {
        std::vector<int32_t> scratch(wave_compile_scratch_ints(len));
        HostWave w;
        n = compile_wave(w, tokens, values, len, nv, wave_compile_scratch(scratch.data(), len), ins.data(), &sn,
                         pf);
    }
    if (n < 0) return SR_ERR_BAD_EXPR;
    for (int i = 0; i < n; ++i) {
        out_code[i] = ins[i].code;
        out_arg[i] = ins[i].arg;
    }
    *stack_need = sn;
    return n;
}

int sr_compile_param_expr(const int32_t *tokens, const float *values, const int32_t *const_slot,
                          int32_t len, int32_t nv, int32_t *out_code, int32_t *out_arg, int32_t *out_aux,
                          int32_t *stack_need, int32_t *tape_rows)
{
    if (!tokens || !values || !const_slot || !out_code || !out_arg || !stack_need) return SR_ERR_INVALID_ARG;
    if (len <= 0 || len > MAX_EXPR_LEN || raw_stack_depth(tokens, len) < 0) return SR_ERR_BAD_EXPR;
    std::vector<int> size(len), need(len), leaf(len), pstack(len);
    std::vector<CompileFrame> frames(len + 1);
    std::vector<Ins> ins(len);
    int sn = 0;
    int n = compile_expr(tokens, values, len, nv, size.data(), need.data(), frames.data(), ins.data(), &sn,
                         leaf.data());
    if (n < 0) return SR_ERR_BAD_EXPR;
    if (param_postpass(ins.data(), n, leaf.data(), const_slot, pstack.data()) != 0) return SR_ERR_INVALID_ARG;
    std::vector<int32_t> aux(n);
    const int rows = adjoint_postpass(ins.data(), n, aux.data());
    if (rows < 0) return SR_ERR_CAPACITY;
    if (tape_rows) *tape_rows = rows;
    for (int i = 0; i < n; ++i) {
        out_code[i] = ins[i].code;
        out_arg[i] = ins[i].arg;
        if (out_aux) out_aux[i] = aux[i];
    }
    *stack_need = sn;
    return n;
}

int sr_get_tuning(sr_ctx *x, sr_tuning *out)
{
    if (!x || !out) return SR_ERR_INVALID_ARG;
    *out = x->c.tune;
    return SR_OK;
}

int sr_set_tuning(sr_ctx *x, const sr_tuning *t)
{
    if (!x || !t) return SR_ERR_INVALID_ARG;
    Ctx &c = x->c;
    if (t->dpt < 1 || t->dpt > 4) return c.fail(SR_ERR_INVALID_ARG, "dpt must be 1, 2, 3 or 4");
    if (t->cache_dpt < 0 || t->cache_dpt > 4) return c.fail(SR_ERR_INVALID_ARG, "cache_dpt must be 0..4");
    if (t->kernel < 0 || t->kernel > 1 || t->cache_kernel < 0 || t->cache_kernel > 1)
        return c.fail(SR_ERR_INVALID_ARG, "kernel / cache_kernel must be 0 (lane) or 1 (vector)");
    if (t->chunk < 64 || (t->n_queues != 1 && t->n_queues != 8) || t->blocks_per_cu < 0)
        return c.fail(SR_ERR_INVALID_ARG, "bad tuning");
    c.tune = *t;
    return SR_OK;
}

}  // extern "C"
