// sr_jit.cpp -- hipRTC backend: one generated module per population.
//
// Replaces the NVRTC path of the reference (cl_gpu_nvrtc.cu:152-548: build a
// source with one straight-line evaluation per expression, compile it at run
// time, launch, report compile ms next to execution ms).  Here every
// expression becomes a __device__ function of straight-line SSA code over the
// op functions of sr_ops.h (embedded verbatim, so each node computes exactly
// what the interpreter computes), and one persistent kernel pulls the same
// (expression, datapoint-chunk) work items as the batched interpreter, with
// the same datapoint-per-lane layout and fused MSE partials, so its MSEs are
// bit-identical to SR_EVAL_BATCHED under the same tuning.  X is read in the
// column-major layout the context holds (the reference's NVRTC kernel indexes
// row-major data uploaded column-major, cl_gpu_nvrtc.cu:172 vs :425-431; that
// quirk is not reproduced).
#include <hip/hiprtc.h>

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "sr_ctx.h"
#include "sr_kernels.h"
#include "sr_launch.h"

namespace sr {

static const char kOpsSrc[] =
#include "sr_ops_src.inc"
    ;

static const char *unary_fn(int op)
{
    switch (op) {
    case OP_SIN: return "f_sin";
    case OP_COS: return "f_cos";
    case OP_TAN: return "f_tan";
    case OP_SINH: return "f_sinh";
    case OP_COSH: return "f_cosh";
    case OP_TANH: return "f_tanh";
    case OP_EXP: return "f_exp";
    case OP_LOG: return "f_log";
    case OP_INV: return "f_inv";
    case OP_ASIN: return "f_asin";
    case OP_ACOS: return "f_acos";
    case OP_ATAN: return "f_atan";
    case OP_LOOSE_LOG: return "f_loose_log";
    case OP_LOOSE_INV: return "f_loose_inv";
    case OP_ABS: return "f_abs";
    case OP_NEG: return "f_neg";
    case OP_SQRT: return "f_sqrt";
    case OP_LOOSE_SQRT: return "f_loose_sqrt";
    case OP_EXP_CLAMPED: return "f_exp_clamped";
    default: return "f_zero";
    }
}

static const char *binary_fn(int op)
{
    switch (op) {
    case OP_ADD: return "f_add";
    case OP_SUB: return "f_sub";
    case OP_MUL: return "f_mul";
    case OP_DIV: return "f_div";
    case OP_POW: return "f_pow";
    case OP_MIN: return "f_min";
    case OP_MAX: return "f_max";
    case OP_LOOSE_DIV: return "f_loose_div";
    case OP_LOOSE_POW: return "f_loose_pow";
    default: return "f_zero2";
    }
}

static void appendf(std::string &s, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
static void appendf(std::string &s, const char *fmt, ...)
{
    char buf[256];
    va_list ap;
    va_start(ap, fmt);
    int n = vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (n >= (int)sizeof buf) {
        std::vector<char> big(n + 1);
        va_start(ap, fmt);
        vsnprintf(big.data(), big.size(), fmt, ap);
        va_end(ap);
        s.append(big.data(), n);
    } else {
        s.append(buf, n);
    }
}

// Straight-line body of one expression: the prefix stream is scanned right to
// left as the reference interpreter does (cl_cpu_eval.cpp:59-90); a binary op
// takes the first popped value as its left operand.
static void emit_expr(std::string &out, int id, const int32_t *tok, const float *val, int L, int nv)
{
    appendf(out, "SR_OPF float e%d(const float *x)\n{\n    using namespace sr_ops;\n", id);
    std::vector<std::string> stk;
    int t = 0;
    char name[64];
    for (int i = L - 1; i >= 0; --i) {
        const int k = tok[i];
        if (k == TOK_CONST) {
            uint32_t bits;
            std::memcpy(&bits, &val[i], 4);
            snprintf(name, sizeof name, "__int_as_float(0x%08x)", bits);
            stk.emplace_back(name);
        } else if (k == TOK_VAR) {
            int v = (int)val[i];
            if (v < nv) snprintf(name, sizeof name, "x[%d]", v);
            else snprintf(name, sizeof name, "0.0f");
            stk.emplace_back(name);
        } else if (is_binary(k)) {
            std::string a = stk.back();
            stk.pop_back();
            std::string b = stk.back();
            stk.pop_back();
            appendf(out, "    const float t%d = %s(%s, %s);\n", t, binary_fn(k), a.c_str(), b.c_str());
            snprintf(name, sizeof name, "t%d", t++);
            stk.emplace_back(name);
        } else {
            std::string a = stk.back();
            stk.pop_back();
            appendf(out, "    const float t%d = %s(%s);\n", t, unary_fn(k), a.c_str());
            snprintf(name, sizeof name, "t%d", t++);
            stk.emplace_back(name);
        }
    }
    appendf(out, "    return %s;\n}\n", stk.back().c_str());
}

static const char kKernelSrc[] = R"SRK(
template <int DPT, typename F>
__device__ __forceinline__ void sr_chunk(F f, const float *__restrict__ X, long long N, int lane,
                                         long long c0, long long c1, double &sum, long long &cnt)
{
    for (long long t0 = c0; t0 < c1; t0 += 64 * DPT) {
        float x[DPT][SR_NVP];
        float y[DPT];
        long long dp[DPT];
#pragma unroll
        for (int d = 0; d < DPT; ++d) {
            dp[d] = t0 + d * 64 + lane;
            const long long dc = dp[d] < N ? dp[d] : N - 1;
#pragma unroll
            for (int k = 0; k < SR_NV; ++k) x[d][k] = X[(long long)k * N + dc];
            y[d] = X[(long long)SR_NV * N + dc];
        }
#pragma unroll
        for (int d = 0; d < DPT; ++d) {
            const float p = f(x[d]);
            const bool ok = dp[d] < c1 && __builtin_isfinite(p);
            const float diff = p - y[d];
            const float sq = diff * diff;
            sum += ok ? (double)sq : 0.0;
            cnt += ok ? 1 : 0;
        }
    }
}

extern "C" __global__ __launch_bounds__(256) void sr_jit_mse(
    const float *__restrict__ X, long long N, const int *__restrict__ order, int n_exprs,
    unsigned *counters, int n_queues, long long slice, long long chunk, int cpq, int P,
    double *psum, long long *pcnt)
{
    const int lane = threadIdx.x & 63;
    const int q = n_queues > 1 ? (int)(blockIdx.x % (unsigned)n_queues) : 0;
    const int items = n_exprs * cpq;
    for (;;) {
        int item = 0;
        if (lane == 0) item = (int)atomicAdd(&counters[q], 1u);
        item = __builtin_amdgcn_readfirstlane(item);
        if (item >= items) break;
        const int rank = item / cpq;
        const int ch = item - rank * cpq;
        const int e = order[rank];
        const long long q0 = (long long)q * slice;
        const long long q1 = q0 + slice < N ? q0 + slice : N;
        const long long c0 = q0 + (long long)ch * chunk;
        const long long c1 = c0 + chunk < q1 ? c0 + chunk : q1;
        double sum = 0.0;
        long long cnt = 0;
        switch (e) {
SR_CASES
        default: break;
        }
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) {
            sum += __shfl_xor(sum, m, 64);
            cnt += __shfl_xor(cnt, m, 64);
        }
        if (lane == 0) {
            psum[(long long)e * P + q * cpq + ch] = sum;
            pcnt[(long long)e * P + q * cpq + ch] = cnt;
        }
    }
}
)SRK";

// Full module source for a population (values already fp32).
static std::string jit_source(const sr_population &pop, int nv, int dpt)
{
    std::string s;
    s.reserve(4096 + (size_t)pop.num_exprs * 2048);
    appendf(s, "#define SR_NV %d\n#define SR_NVP %d\n#define SR_DPT %d\n", nv, nv > 0 ? nv : 1, dpt);
    s += "#define SR_OPF_T __device__ __attribute__((noinline))\n";
    s += kOpsSrc;
    std::vector<float> vals;
    for (int e = 0; e < pop.num_exprs; ++e) {
        expr_values_f32(pop, e, vals);
        emit_expr(s, e, pop.tokens + pop.offsets[e], vals.data(), pop.lengths[e], nv);
    }
    std::string cases;
    for (int e = 0; e < pop.num_exprs; ++e)
        appendf(cases, "        case %d: sr_chunk<SR_DPT>(e%d, X, N, lane, c0, c1, sum, cnt); break;\n", e, e);
    std::string k = kKernelSrc;
    const size_t at = k.find("SR_CASES");
    k.replace(at, 8, cases);
    s += k;
    return s;
}

static uint64_t fnv(uint64_t h, const void *p, size_t n)
{
    const unsigned char *b = (const unsigned char *)p;
    for (size_t i = 0; i < n; ++i) {
        h ^= b[i];
        h *= 1099511628211ull;
    }
    return h;
}

// Compile a source for gfx950; returns SR_OK and the code object.
static int rtc_compile(const std::string &src, std::vector<char> &code, std::string &log, double &ms)
{
    auto t0 = std::chrono::steady_clock::now();
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "sr_jit.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
        log = "hiprtcCreateProgram failed";
        return SR_ERR_JIT;
    }
    const char *opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
                          // packed FP32 (v_pk_*_f32) stays off, as in the prebuilt kernels
                          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    hiprtcResult r = hiprtcCompileProgram(prog, (int)(sizeof opts / sizeof opts[0]), opts);
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    log.assign(ls, '\0');
    if (ls) hiprtcGetProgramLog(prog, &log[0]);
    if (r != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&prog);
        log = std::string("hiprtc: ") + hiprtcGetErrorString(r) + "\n" + log;
        return SR_ERR_JIT;
    }
    size_t cs = 0;
    hiprtcGetCodeSize(prog, &cs);
    code.resize(cs);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return SR_OK;
}

JitState::~JitState()
{
    if (mod) hipModuleUnload(mod);
}

// Device execution of generated modules is switched off: the first GPU run of
// this path (100 expressions, 10 variables, 100k datapoints) stopped making
// progress after the compile, and again after the code-object lifetime fix;
// the cause is not identified yet.  Source generation and host compilation
// (sr_jit_source / sr_jit_compile / sr_jit_code) stay available.
static constexpr bool kJitDeviceEnabled = false;

int eval_jit(Ctx &c, const sr_population &pop, sr_stats *st)
{
    if (!kJitDeviceEnabled)
        return c.fail(SR_ERR_UNSUPPORTED,
                      "SR_EVAL_JIT: device execution of hipRTC modules is disabled in this build "
                      "(host-side sr_jit_source / sr_jit_compile work)");
    const int E = pop.num_exprs;
    Program check;
    int rc = build_program(c, pop, false, check);  // validation + lengths for the order
    if (rc) return rc;
    LaunchPlan lp;
    rc = plan_batched(c, MODE_MSE, E, lp);
    if (rc) return rc;
    JitState &j = c.jit;
    uint64_t key = 1469598103934665603ull;
    const int32_t hdr[3] = {c.nv, lp.dpt, E};
    key = fnv(key, hdr, sizeof hdr);
    std::vector<float> vals;
    for (int e = 0; e < E; ++e) {
        key = fnv(key, pop.tokens + pop.offsets[e], (size_t)pop.lengths[e] * 4);
        expr_values_f32(pop, e, vals);
        key = fnv(key, vals.data(), vals.size() * 4);
    }
    double compile_ms = 0.0;
    hipError_t he;
    if (!j.mod || key != j.key) {
        if (j.mod) {
            hipModuleUnload(j.mod);
            j.mod = nullptr;
        }
        std::string log;
        j.image.clear();
        rc = rtc_compile(jit_source(pop, c.nv, lp.dpt), j.image, log, compile_ms);
        if (rc) return c.fail(rc, log);
        if ((he = hipModuleLoadData(&j.mod, j.image.data())) != hipSuccess) return c.hip_fail(he, "hipModuleLoadData");
        if ((he = hipModuleGetFunction(&j.fn, j.mod, "sr_jit_mse")) != hipSuccess)
            return c.hip_fail(he, "hipModuleGetFunction");
        int occ = 0;
        if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&occ, j.fn, kBlock, 0) != hipSuccess || occ <= 0)
            occ = 1;
        j.occ = occ;
        j.key = key;
        j.code_bytes = (int64_t)j.image.size();
        j.last_compile_ms = compile_ms;
    }
    // order + partial buffers
    if ((he = c.d_meta.ensure((size_t)E * 4)) != hipSuccess) return c.hip_fail(he, "hipMalloc");
    if ((he = c.h_meta.ensure((size_t)E * 4)) != hipSuccess) return c.hip_fail(he, "hipHostMalloc");
    std::memcpy(c.h_meta.p, check.order.data(), (size_t)E * 4);
    hipStream_t s = c.stream;
    hipEventRecord(c.ev[0], s);
    if ((he = hipMemcpyAsync(c.d_meta.p, c.h_meta.p, (size_t)E * 4, hipMemcpyHostToDevice, s)) != hipSuccess)
        return c.hip_fail(he, "H2D order");
    hipEventRecord(c.ev[1], s);
    uint32_t *counters = c.d_counters.as<uint32_t>();
    if ((he = launch_zero_u32(counters, lp.nq, s)) != hipSuccess) return c.hip_fail(he, "zero counters");
    const float *X = c.X;
    long long N = c.N, slice = lp.slice, chunk = lp.chunk;
    const int *order = c.d_meta.as<int32_t>();
    int n_exprs = E, nq = lp.nq, cpq = lp.cpq, P = lp.P;
    double *psum = c.d_psum.as<double>();
    long long *pcnt = c.d_pcnt.as<long long>();
    void *args[] = {&X, &N, &order, &n_exprs, &counters, &nq, &slice, &chunk, &cpq, &P, &psum, &pcnt};
    const int grid = (int)(((int64_t)c.num_cu * j.occ + nq - 1) / nq * nq);
    hipEventRecord(c.ev[4], s);
    if ((he = hipModuleLaunchKernel(j.fn, grid, 1, 1, kBlock, 1, 1, 0, s, args, nullptr)) != hipSuccess)
        return c.hip_fail(he, "JIT launch");
    hipEventRecord(c.ev[5], s);
    if ((he = launch_finalize(psum, pcnt, P, E, c.d_mse.as<float>(), c.d_cnt.as<long long>(), s)) != hipSuccess)
        return c.hip_fail(he, "finalize launch");
    hipEventRecord(c.ev[2], s);
    if (st) {
        st->jit_compile_time_ms += compile_ms;
        st->num_kernel_launches += 3;
        st->executed_node_evals += check.tokens * c.N;
        if ((he = hipEventSynchronize(c.ev[5])) != hipSuccess) return c.hip_fail(he, "JIT kernel");
        float ms = 0.f;
        hipEventElapsedTime(&ms, c.ev[4], c.ev[5]);
        st->interp_kernel_ms += ms;
        st->interp_launches += 1;
    }
    return SR_OK;
}

}  // namespace sr

using namespace sr;

extern "C" {

int sr_jit_source(const sr_population *pop, int32_t nv, int32_t dpt, char *out, int64_t cap, int64_t *len)
{
    if (!pop || !len || nv < 0 || nv > MAX_VARS || (dpt != 1 && dpt != 2 && dpt != 4)) return SR_ERR_INVALID_ARG;
    std::vector<float> vals;
    for (int e = 0; e < pop->num_exprs; ++e) {
        const int32_t *tok = pop->tokens + pop->offsets[e];
        if (pop->lengths[e] <= 0 || pop->lengths[e] > MAX_EXPR_LEN || raw_stack_depth(tok, pop->lengths[e]) < 0)
            return SR_ERR_BAD_EXPR;
        expr_values_f32(*pop, e, vals);
        for (int k = 0; k < pop->lengths[e]; ++k)
            if (tok[k] == OP_REF || (tok[k] == TOK_VAR && !(vals[k] >= 0.0f && vals[k] < 1e9f)))
                return SR_ERR_BAD_EXPR;
    }
    std::string s = jit_source(*pop, nv, dpt);
    *len = (int64_t)s.size();
    if (out && cap > 0) {
        const size_t n = std::min<size_t>(s.size(), (size_t)cap - 1);
        std::memcpy(out, s.data(), n);
        out[n] = '\0';
    }
    return SR_OK;
}

int sr_jit_code(const sr_population *pop, int32_t nv, int32_t dpt, void *out, int64_t cap, int64_t *len)
{
    int64_t slen = 0;
    int rc = sr_jit_source(pop, nv, dpt, nullptr, 0, &slen);
    if (rc) return rc;
    if (!len) return SR_ERR_INVALID_ARG;
    std::string log;
    std::vector<char> code;
    double ms = 0.0;
    rc = rtc_compile(jit_source(*pop, nv, dpt), code, log, ms);
    if (rc) return rc;
    *len = (int64_t)code.size();
    if (out && cap >= (int64_t)code.size()) std::memcpy(out, code.data(), code.size());
    return SR_OK;
}

int sr_jit_compile(const sr_population *pop, int32_t nv, int32_t dpt, char *log_out, int64_t log_cap,
                   double *compile_ms, int64_t *code_bytes)
{
    int64_t len = 0;
    int rc = sr_jit_source(pop, nv, dpt, nullptr, 0, &len);
    if (rc) return rc;
    std::string log;
    std::vector<char> code;
    double ms = 0.0;
    rc = rtc_compile(jit_source(*pop, nv, dpt), code, log, ms);
    if (log_out && log_cap > 0) {
        const size_t n = std::min<size_t>(log.size(), (size_t)log_cap - 1);
        std::memcpy(log_out, log.data(), n);
        log_out[n] = '\0';
    }
    if (compile_ms) *compile_ms = ms;
    if (code_bytes) *code_bytes = (int64_t)code.size();
    return rc;
}

}  // extern "C"
