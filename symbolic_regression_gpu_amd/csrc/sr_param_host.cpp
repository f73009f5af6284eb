// sr_param_host.cpp -- C ABI of constant fitting: parametric programs whose
// free constants live in a device array [E][8], evaluated forward (MSE),
// forward on 2*8 central-difference replicas, or forward + adjoint (MSE and
// dMSE/dconstants).  All entry points take device pointers and only enqueue
// work on the context's stream, so a torch.autograd.Function can call them on
// torch's current stream.
#include <algorithm>
#include <cstring>
#include <numeric>

#include "sr_adjoint.h"
#include "sr_adjoint_prog.h"
#include "sr_ctx.h"
#include "sr_launch.h"
#include "sr_param_launch.h"

namespace sr {

static int64_t rup(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

struct ParamPlan {
    int dpt, xc, nq, cpq, P, grid;
    int64_t slice, chunk;
};

static int param_plan(Ctx &c, bool adjoint, ParamPlan &pp, int force_dpt = 0)
{
    // variable classes of the constant-fitting kernels: <= 4, <= 8 (nine
    // registers, GPR-indexed), <= 16, <= 32
    pp.xc = c.nv <= 8 && c.nv > 4 ? 8 : pick_xc(c.nv);
    const int max_dpt = adjoint ? (pp.xc == 32 ? 1 : 2) : (pp.xc == 32 ? 1 : 4);
    const int want = force_dpt ? force_dpt : (adjoint ? c.par.adj_dpt : c.par.fwd_dpt);
    if (force_dpt && force_dpt > max_dpt) return SR_ERR_UNSUPPORTED;
    pp.dpt = std::max(1, std::min(want ? want : c.tune.dpt, max_dpt));
    if (pp.dpt == 3) pp.dpt = 2;
    pp.nq = c.tune.n_queues > 1 ? c.tune.n_queues : 1;
    const int64_t tile = 64 * pp.dpt;
    pp.slice = rup((c.N + pp.nq - 1) / pp.nq, 512);
    pp.chunk = rup(std::max<int64_t>(c.tune.chunk, tile), tile);
    if (pp.chunk > pp.slice) pp.chunk = pp.slice;
    pp.cpq = (int)((pp.slice + pp.chunk - 1) / pp.chunk);
    pp.P = pp.nq * pp.cpq;
    const int prog_lds = c.par.max_len + 1;
    int occ = adjoint ? adjoint_occupancy(pp.dpt, pp.xc, prog_lds) : param_fwd_occupancy(pp.dpt, pp.xc, prog_lds);
    if (occ <= 0) return c.fail(SR_ERR_UNSUPPORTED, "no constant-fitting kernel for this variable count");
    pp.grid = (int)rup((int64_t)c.num_cu * occ, pp.nq);
    return SR_OK;
}

static int param_ready(Ctx &c)
{
    if (!c.par.ready) return c.fail(SR_ERR_INVALID_ARG, "no parametric population: call sr_param_prepare");
    if (!c.X) return c.fail(SR_ERR_NO_DATA, "no data: call sr_set_data first");
    return SR_OK;
}

static ParamArgs base_args(Ctx &c, const ParamPlan &pp, const float *cvals)
{
    ParamState &s = c.par;
    ParamArgs a{};
    a.progs = s.progs.as<Ins>();
    a.aux = s.aux.as<int32_t>();
    a.prog_off = s.meta.as<int32_t>();
    a.prog_len = a.prog_off + s.E;
    a.order = a.prog_off + 2 * s.E;
    a.nslots = a.prog_off + 3 * s.E;
    a.X = c.X;
    a.cvals = cvals;
    a.counters = c.d_counters.as<uint32_t>();
    a.N = c.N;
    a.slice = pp.slice;
    a.chunk = pp.chunk;
    a.n_exprs = s.E;
    a.nv = c.nv;
    a.n_queues = pp.nq;
    a.chunks_per_queue = pp.cpq;
    a.P = pp.P;
    a.replicas = 1;
    a.prog_lds = s.max_len + 1;
    return a;
}

// Forward MSE of `replicas` constant sets into mse_out [replicas][E].
static int param_forward(Ctx &c, const float *cvals, int replicas, float h, float *mse_out)
{
    ParamState &s = c.par;
    ParamPlan pp;
    int rc = param_plan(c, false, pp);
    if (rc) return rc;
    hipError_t he;
    const size_t rows = (size_t)replicas * s.E;
    if ((he = s.psum.ensure(rows * pp.P * 8)) != hipSuccess) return c.hip_fail(he, "hipMalloc");
    if ((he = s.pcnt.ensure(rows * pp.P * 8)) != hipSuccess) return c.hip_fail(he, "hipMalloc");
    if ((he = c.d_counters.ensure(256)) != hipSuccess) return c.hip_fail(he, "hipMalloc");
    ParamArgs a = base_args(c, pp, cvals);
    a.psum = s.psum.as<double>();
    a.pcnt = s.pcnt.as<long long>();
    a.replicas = replicas;
    a.fd_h = h;
    hipStream_t st = c.stream;
    if ((he = launch_zero_u32(a.counters, pp.nq, st)) != hipSuccess) return c.hip_fail(he, "zero");
    if ((he = launch_param_fwd(pp.dpt, pp.xc, a, pp.grid, st)) != hipSuccess)
        return c.hip_fail(he, "parametric forward");
    if ((he = launch_finalize(a.psum, a.pcnt, pp.P, (int)rows, mse_out, nullptr, st)) != hipSuccess)
        return c.hip_fail(he, "finalize");
    return SR_OK;
}

static int param_adjoint(Ctx &c, const float *d_cvals, float *d_mse, float *d_grad, int force_dpt)
{
    ParamState &s = c.par;
    ParamPlan pp;
    int rc = param_plan(c, true, pp, force_dpt);
    if (rc) return rc;
    hipError_t he;
    const size_t waves = (size_t)pp.grid * (kBlock / kWave);
    if ((he = s.psum.ensure((size_t)s.E * pp.P * 8)) != hipSuccess) return c.hip_fail(he, "hipMalloc");
    if ((he = s.pcnt.ensure((size_t)s.E * pp.P * 8)) != hipSuccess) return c.hip_fail(he, "hipMalloc");
    if ((he = s.pgrad.ensure((size_t)s.E * pp.P * MAX_SLOTS * 8)) != hipSuccess) return c.hip_fail(he, "hipMalloc");
    if ((he = s.tape.ensure(waves * (size_t)s.tape_rows * kWave * pp.dpt * 4)) != hipSuccess)
        return c.hip_fail(he, "hipMalloc tape");
    if ((he = c.d_counters.ensure(256)) != hipSuccess) return c.hip_fail(he, "hipMalloc");
    ParamArgs a = base_args(c, pp, d_cvals);
    a.psum = s.psum.as<double>();
    a.pcnt = s.pcnt.as<long long>();
    a.pgrad = s.pgrad.as<double>();
    a.tape = s.tape.as<float>();
    a.tape_stride = s.tape_rows;
    hipStream_t st = c.stream;
    if ((he = launch_zero_u32(a.counters, pp.nq, st)) != hipSuccess) return c.hip_fail(he, "zero");
    if ((he = launch_adjoint(pp.dpt, pp.xc, a, pp.grid, st)) != hipSuccess) return c.hip_fail(he, "adjoint");
    if ((he = launch_finalize_grad(a.psum, a.pcnt, a.pgrad, pp.P, s.E, d_mse, d_grad, st)) != hipSuccess)
        return c.hip_fail(he, "finalize grad");
    return SR_OK;
}

// Time one adjoint or one forward pass (replicas = 1) of the prepared
// population at a given datapoints-per-lane (the autotuner's probe).
int sr_param_time(Ctx &c, bool adjoint, int dpt, const float *d_cvals, float *ms)
{
    int rc = param_ready(c);
    if (rc) return rc;
    ParamState &s = c.par;
    hipError_t he;
    DevBuf tmp;
    if ((he = tmp.ensure((size_t)s.E * (MAX_SLOTS + 1) * 4)) != hipSuccess) return c.hip_fail(he, "hipMalloc");
    float *mse = tmp.as<float>(), *grad = mse + s.E;
    const int saved_f = s.fwd_dpt, saved_a = s.adj_dpt;
    for (int rep = 0; rep < 2; ++rep) {
        hipEventRecord(c.ev[4], c.stream);
        if (adjoint) {
            rc = param_adjoint(c, d_cvals, mse, grad, dpt);
        } else {
            ParamPlan probe;
            rc = param_plan(c, false, probe, dpt);
            if (!rc) {
                s.fwd_dpt = dpt;
                rc = param_forward(c, d_cvals, 1, 0.0f, mse);
                s.fwd_dpt = saved_f;
            }
        }
        hipEventRecord(c.ev[5], c.stream);
        if (rc) break;
    }
    s.adj_dpt = saved_a;
    if (rc) return rc;
    if ((he = hipEventSynchronize(c.ev[5])) != hipSuccess) return c.hip_fail(he, "param timing");
    hipEventElapsedTime(ms, c.ev[4], c.ev[5]);
    return SR_OK;
}

}  // namespace sr

using namespace sr;

extern "C" {

int sr_param_prepare(sr_ctx *x, const sr_population *pop, const int32_t *const_slot, float *init_vals)
{
    if (!x || !pop || !const_slot) return SR_ERR_INVALID_ARG;
    Ctx &c = x->c;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    hipSetDevice(c.device);
    if (!c.X) return c.fail(SR_ERR_NO_DATA, "no data: call sr_set_data first");
    const int E = pop->num_exprs;
    if (E <= 0) return c.fail(SR_ERR_INVALID_ARG, "empty population");
    Program check;
    int rc = build_program(c, *pop, true, check);
    if (rc) return rc;
    ParamState &s = c.par;
    std::vector<Ins> ins;
    std::vector<int32_t> aux_all, aux_tmp(MAX_EXPR_LEN);
    std::vector<int32_t> off(E), len(E), nslots(E), order(E);
    int tape_rows = 1;
    std::vector<float> vals, cv((size_t)E * MAX_SLOTS, 0.0f);
    std::vector<int> size(MAX_EXPR_LEN), need(MAX_EXPR_LEN), leaf(MAX_EXPR_LEN), pstack(MAX_EXPR_LEN);
    std::vector<CompileFrame> frames(MAX_EXPR_LEN + 1);
    std::vector<Ins> tmp(MAX_EXPR_LEN);
    int max_len = 1;
    for (int e = 0; e < E; ++e) {
        const int32_t L = pop->lengths[e];
        const int32_t *tok = pop->tokens + pop->offsets[e];
        const int32_t *slot = const_slot + pop->offsets[e];
        expr_values_f32(*pop, e, vals);
        int used = 0;
        std::vector<bool> seen(MAX_SLOTS, false);
        for (int k = 0; k < L; ++k) {
            if (slot[k] < 0) continue;
            if (tok[k] != TOK_CONST || slot[k] >= MAX_SLOTS)
                return c.fail(SR_ERR_INVALID_ARG, "expression " + std::to_string(e) +
                                                      ": free slot on a non-constant token or slot >= 8");
            used = std::max(used, slot[k] + 1);
            if (!seen[slot[k]]) {
                seen[slot[k]] = true;
                cv[(size_t)e * MAX_SLOTS + slot[k]] = vals[k];
            }
        }
        int sn = 0;
        int n = compile_expr(tok, vals.data(), L, c.nv, size.data(), need.data(), frames.data(), tmp.data(),
                             &sn, leaf.data());
        if (n < 0 || sn > kStackSlots) return c.fail(SR_ERR_BAD_EXPR, "expression " + std::to_string(e));
        if (param_postpass(tmp.data(), n, leaf.data(), slot, pstack.data()) != 0)
            return c.fail(SR_ERR_INVALID_ARG, "bad free-constant slot");
        const int rows = adjoint_postpass(tmp.data(), n, aux_tmp.data());
        if (rows < 0) return c.fail(SR_ERR_CAPACITY, "expression " + std::to_string(e) + ": adjoint tape too long");
        tape_rows = std::max(tape_rows, rows);
        aux_all.insert(aux_all.end(), aux_tmp.begin(), aux_tmp.begin() + n);
        off[e] = (int32_t)ins.size();
        len[e] = n;
        nslots[e] = used;
        max_len = std::max(max_len, n);
        ins.insert(ins.end(), tmp.begin(), tmp.begin() + n);
    }
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return len[a] > len[b]; });
    hipError_t he;
    if ((he = s.progs.ensure(ins.size() * sizeof(Ins))) != hipSuccess) return c.hip_fail(he, "hipMalloc");
    if ((he = s.aux.ensure(aux_all.size() * 4)) != hipSuccess) return c.hip_fail(he, "hipMalloc");
    if ((he = s.meta.ensure((size_t)E * 16)) != hipSuccess) return c.hip_fail(he, "hipMalloc");
    std::vector<int32_t> meta((size_t)E * 4);
    std::memcpy(meta.data(), off.data(), E * 4);
    std::memcpy(meta.data() + E, len.data(), E * 4);
    std::memcpy(meta.data() + 2 * E, order.data(), E * 4);
    std::memcpy(meta.data() + 3 * E, nslots.data(), E * 4);
    he = hipMemcpyAsync(s.progs.p, ins.data(), ins.size() * sizeof(Ins), hipMemcpyHostToDevice, c.stream);
    if (he == hipSuccess) he = hipMemcpyAsync(s.meta.p, meta.data(), meta.size() * 4, hipMemcpyHostToDevice, c.stream);
    if (he == hipSuccess)
        he = hipMemcpyAsync(s.aux.p, aux_all.data(), aux_all.size() * 4, hipMemcpyHostToDevice, c.stream);
    if (he == hipSuccess) he = hipStreamSynchronize(c.stream);
    if (he != hipSuccess) return c.hip_fail(he, "H2D parametric programs");
    s.E = E;
    s.max_len = max_len;
    s.tape_rows = tape_rows;
    s.nslots = nslots;
    s.tokens = check.tokens;
    s.ready = true;
    if (init_vals) std::memcpy(init_vals, cv.data(), cv.size() * 4);
    return SR_OK;
}

int sr_param_eval(sr_ctx *x, const float *d_cvals, float *d_mse)
{
    if (!x || !d_cvals || !d_mse) return SR_ERR_INVALID_ARG;
    Ctx &c = x->c;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    hipSetDevice(c.device);
    int rc = param_ready(c);
    if (rc) return rc;
    return param_forward(c, d_cvals, 1, 0.0f, d_mse);
}

int sr_param_fd_grad(sr_ctx *x, const float *d_cvals, float h, float *d_grad)
{
    if (!x || !d_cvals || !d_grad || !(h > 0.0f)) return SR_ERR_INVALID_ARG;
    Ctx &c = x->c;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    hipSetDevice(c.device);
    int rc = param_ready(c);
    if (rc) return rc;
    ParamState &s = c.par;
    hipError_t he = s.fd_mse.ensure((size_t)2 * MAX_SLOTS * s.E * 4);
    if (he != hipSuccess) return c.hip_fail(he, "hipMalloc");
    rc = param_forward(c, d_cvals, 2 * MAX_SLOTS, h, s.fd_mse.as<float>());
    if (rc) return rc;
    he = launch_fd_combine(d_cvals, s.fd_mse.as<float>(), s.meta.as<int32_t>() + 3 * s.E, s.E, h, d_grad,
                           c.stream);
    if (he != hipSuccess) return c.hip_fail(he, "fd combine");
    return SR_OK;
}

int sr_param_grad(sr_ctx *x, const float *d_cvals, float *d_mse, float *d_grad)
{
    if (!x || !d_cvals || !d_mse || !d_grad) return SR_ERR_INVALID_ARG;
    Ctx &c = x->c;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    hipSetDevice(c.device);
    int rc = param_ready(c);
    if (rc) return rc;
    return param_adjoint(c, d_cvals, d_mse, d_grad, 0);
}

int sr_param_info(sr_ctx *x, int32_t *num_exprs, int32_t *nslots_out, int64_t *tokens)
{
    if (!x) return SR_ERR_INVALID_ARG;
    Ctx &c = x->c;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    if (!c.par.ready) return c.fail(SR_ERR_INVALID_ARG, "no parametric population");
    if (num_exprs) *num_exprs = c.par.E;
    if (nslots_out) std::memcpy(nslots_out, c.par.nslots.data(), c.par.nslots.size() * 4);
    if (tokens) *tokens = c.par.tokens;
    return SR_OK;
}

int sr_grad_consts(sr_ctx *x, const sr_population *pop, const uint8_t *const_mask, float *grad_out,
                   float *mse_out)
{
    if (!x || !pop || !const_mask || !grad_out || pop->num_exprs <= 0 || !pop->tokens || !pop->offsets ||
        !pop->lengths)
        return SR_ERR_INVALID_ARG;
    Ctx &c = x->c;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    hipSetDevice(c.device);
    const int E = pop->num_exprs;
    int64_t T = 0;
    for (int e = 0; e < E; ++e) T = std::max<int64_t>(T, (int64_t)pop->offsets[e] + pop->lengths[e]);
    // one parameter slot per masked constant, in token order
    std::vector<int32_t> slot((size_t)T, -1);
    for (int e = 0; e < E; ++e) {
        int k = 0;
        for (int32_t t = 0; t < pop->lengths[e]; ++t) {
            const int64_t q = (int64_t)pop->offsets[e] + t;
            if (!const_mask[q]) continue;
            if (pop->tokens[q] != TOK_CONST)
                return c.fail(SR_ERR_INVALID_ARG, "expression " + std::to_string(e) + ": mask on a non-constant token");
            if (k >= MAX_SLOTS)
                return c.fail(SR_ERR_UNSUPPORTED, "expression " + std::to_string(e) + ": more than 8 masked constants");
            slot[q] = k++;
        }
    }
    std::vector<float> cv((size_t)E * MAX_SLOTS, 0.0f), mse(E), g((size_t)E * MAX_SLOTS);
    int rc = sr_param_prepare(x, pop, slot.data(), cv.data());
    if (rc) return rc;
    DevBuf d_c, d_m, d_g;
    hipError_t he = d_c.ensure(cv.size() * 4);
    if (he == hipSuccess) he = d_m.ensure((size_t)E * 4);
    if (he == hipSuccess) he = d_g.ensure(g.size() * 4);
    if (he == hipSuccess) he = hipMemcpyAsync(d_c.p, cv.data(), cv.size() * 4, hipMemcpyHostToDevice, c.stream);
    if (he != hipSuccess) return c.hip_fail(he, "gradient buffers");
    if ((rc = sr_param_grad(x, d_c.as<float>(), d_m.as<float>(), d_g.as<float>()))) return rc;
    he = hipMemcpyAsync(g.data(), d_g.p, g.size() * 4, hipMemcpyDeviceToHost, c.stream);
    if (he == hipSuccess) he = hipMemcpyAsync(mse.data(), d_m.p, (size_t)E * 4, hipMemcpyDeviceToHost, c.stream);
    if (he == hipSuccess) he = hipStreamSynchronize(c.stream);
    if (he != hipSuccess) return c.hip_fail(he, "gradient readback");
    for (int e = 0; e < E; ++e)
        for (int32_t t = 0; t < pop->lengths[e]; ++t) {
            const int64_t q = (int64_t)pop->offsets[e] + t;
            grad_out[q] = slot[q] >= 0 ? g[(size_t)e * MAX_SLOTS + slot[q]] : 0.0f;
        }
    if (mse_out) std::memcpy(mse_out, mse.data(), (size_t)E * 4);
    return SR_OK;
}

int sr_sync(sr_ctx *x)
{
    if (!x) return SR_ERR_INVALID_ARG;
    Ctx &c = x->c;
    hipError_t he = hipStreamSynchronize(c.stream);
    if (he != hipSuccess) return c.hip_fail(he, "sync");
    return SR_OK;
}

}  // extern "C"
