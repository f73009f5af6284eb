// sr_noalloc.hip -- sr_eval_mse_noalloc: batched MSE evaluation of a
// device-resident packed population that allocates nothing.
//
// For callers that cannot allocate (code under HIP graph capture, serving
// loops): every intermediate buffer of the batched path -- bytecode programs,
// compile scratch, program order, MSE partials, work counters -- lives in one
// file-scope device array, `scratch`, of kScratchBytes per device.  The
// population is compiled on the device (the same compile_expr the host path
// runs), then the same k_batched / k_finalize kernels run with the same
// launch plan as sr_eval_mse in SR_EVAL_BATCHED, so MSEs and valid counts are
// bit-identical to it.
//
// Every call on a device shares that device's array and leaves its contents
// behind; what must start at zero (work counters, status word) is cleared by
// kernels on the stream at the start of each call.  Calls that overlap on one
// device (two streams, two contexts, two threads) race on the array.
#include <mutex>

#include "sr_ctx.h"
#include "sr_ga_launch.h"
#include "sr_kernels.h"
#include "sr_launch.h"

namespace sr {

constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on each device (each device holds its own
// copy of the array).  Host code uses these addresses, never the array's
// name, which on the host is only a stand-in for the symbol.
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices];

// Called with `device` current (context creation).
hipError_t noalloc_scratch_init(int device, void **out)
{
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

// Status bits of one call (device word in the scratch; not read back).
constexpr int32_t kNaBadExpr = 1, kNaCapacity = 2, kNaStack = 4;

// Exclusive prefix sum of the valid expression lengths: program e goes to
// prog_off[e]; an expression whose program would pass cap_ins gets -1.
__global__ __launch_bounds__(1024) void k_noalloc_scan(const int32_t *len, int E, int64_t cap_ins,
                                                       int32_t *prog_off)
{
    __shared__ long long part[1024];
    const int t = threadIdx.x;
    const int per = (E + 1023) / 1024;
    const int a = min(E, t * per), b = min(E, a + per);
    auto valid_len = [](int32_t l) { return (l >= 1 && l <= MAX_EXPR_LEN) ? (long long)l : 0ll; };
    long long s = 0;
    for (int i = a; i < b; ++i) s += valid_len(len[i]);
    part[t] = s;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        const long long v = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    long long base = t ? part[t - 1] : 0;
    for (int i = a; i < b; ++i) {
        const long long l = valid_len(len[i]);
        prog_off[i] = base + l <= cap_ins ? (int32_t)base : -1;
        base += l;
    }
}

// One thread per expression: validate as build_program does (length, stream
// shape, variable indices, no cache references), then compile to bytecode at
// prog_off[e].  A rejected expression runs the one-instruction program
// "LOAD NaN" kept at progs[nan_at] (MSE NaN, valid count 0).
__global__ void k_noalloc_compile(const int32_t *tok, const float *val, const int32_t *off, const int32_t *len,
                                  int E, int64_t num_tokens, int nv, int32_t *prog_off, Ins *progs,
                                  int32_t *prog_len, int32_t *size_s, int32_t *need_s, CompileFrame *frames,
                                  int32_t nan_at, int32_t *status)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e == 0) {
        progs[nan_at].code = 0 | (SRC_CONST << 8);
        progs[nan_at].arg = f32_bits(__builtin_nanf(""));
    }
    if (e >= E) return;
    const int L = len[e];
    const int64_t o = off[e];
    const int po = prog_off[e];
    bool ok = L >= 1 && L <= MAX_EXPR_LEN && o >= 0 && o + L <= num_tokens;
    int32_t why = ok && po < 0 ? kNaCapacity : kNaBadExpr;
    ok = ok && po >= 0;
    const int32_t *t = tok + (ok ? o : 0);
    const float *v = val + (ok ? o : 0);
    if (ok) ok = raw_stack_depth(t, L) >= 0;
    for (int k = 0; ok && k < L; ++k)
        if ((t[k] == TOK_VAR && !(v[k] >= 0.0f)) || t[k] == OP_REF) ok = false;
    int n = -1, sn = 0;
    if (ok) {
        n = compile_expr(t, v, L, nv, size_s + po, need_s + po, frames + po + e, progs + po, &sn);
        if (n >= 0 && sn > kStackSlots) why = kNaStack;
    }
    if (n < 0 || sn > kStackSlots) {
        atomicOr(status, why);
        prog_off[e] = nan_at;
        n = 1;
    }
    prog_len[e] = n;
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace sr

using namespace sr;

extern "C" int sr_eval_mse_noalloc(sr_ctx *x, int32_t num_exprs, int64_t num_tokens, const int32_t *d_tokens,
                                   const float *d_values, const int32_t *d_offsets, const int32_t *d_lengths,
                                   float *d_mse, int64_t *d_valid)
{
    if (!x) return SR_ERR_INVALID_ARG;
    Ctx &c = x->c;
    std::lock_guard<std::recursive_mutex> g(c.mu);
    if (num_exprs < 0 || num_tokens < 0 || !d_mse ||
        (num_exprs > 0 && (!d_tokens || !d_values || !d_offsets || !d_lengths)))
        return c.fail(SR_ERR_INVALID_ARG, "sr_eval_mse_noalloc: missing or negative argument");
    if (!c.X) return c.fail(SR_ERR_NO_DATA, "no data: call sr_set_data first");
    if (!c.noalloc_scratch) return c.fail(SR_ERR_HIP, "sr_eval_mse_noalloc: no static scratch on this device");
    if (num_exprs == 0) return SR_OK;
    hipSetDevice(c.device);
    LaunchPlan lp;
    int rc = plan_launch(c, MODE_MSE, lp, false);  // the lane interpreter: no spill area
    if (rc) return rc;
    // Scratch layout for this call (every region 256-byte aligned):
    //   programs (num_tokens + 1 instructions, the last one LOAD NaN),
    //   compile scratch (sizes, needs: num_tokens ints; frames: num_tokens +
    //   num_exprs), program offsets / lengths / order (num_exprs ints),
    //   MSE partials (num_exprs * P doubles + counts), counters, status.
    const size_t E = (size_t)num_exprs, T = (size_t)num_tokens, P = (size_t)lp.P;
    const size_t sizes[] = {align256((T + 1) * sizeof(Ins)),     align256(T * 4),     align256(T * 4),
                            align256((T + E) * sizeof(CompileFrame)), align256(E * 4), align256(E * 4),
                            align256(E * 4),                     align256(E * P * 8), align256(E * P * 8),
                            align256(64 * 4),                    align256(4)};
    size_t need = 0;
    for (size_t b : sizes) need += b;
    if (T >= (size_t)INT32_MAX || need > kScratchBytes)
        return c.fail(SR_ERR_CAPACITY,
                      "sr_eval_mse_noalloc: " + std::to_string(num_exprs) + " expressions, " +
                          std::to_string(num_tokens) + " tokens and " + std::to_string(lp.P) +
                          " partials per expression need " + std::to_string(need) +
                          " bytes of scratch; the static scratch holds " + std::to_string(kScratchBytes));
    unsigned char *base = static_cast<unsigned char *>(c.noalloc_scratch);
    size_t at = 0;
    auto take = [&](int k) {
        void *p = base + at;
        at += sizes[k];
        return p;
    };
    Ins *progs = (Ins *)take(0);
    int32_t *size_s = (int32_t *)take(1);
    int32_t *need_s = (int32_t *)take(2);
    CompileFrame *frames = (CompileFrame *)take(3);
    int32_t *prog_off = (int32_t *)take(4);
    int32_t *prog_len = (int32_t *)take(5);
    int32_t *order = (int32_t *)take(6);
    double *psum = (double *)take(7);
    long long *pcnt = (long long *)take(8);
    uint32_t *counters = (uint32_t *)take(9);
    int32_t *status = (int32_t *)take(10);

    hipStream_t s = c.stream;
    hipError_t he = launch_zero_u32((uint32_t *)status, 1, s);
    if (he == hipSuccess) {
        hipLaunchKernelGGL(k_noalloc_scan, dim3(1), dim3(1024), 0, s, d_lengths, num_exprs, (int64_t)T, prog_off);
        he = hipGetLastError();
    }
    if (he == hipSuccess) {
        hipLaunchKernelGGL(k_noalloc_compile, dim3((num_exprs + 63) / 64), dim3(64), 0, s, d_tokens, d_values,
                           d_offsets, d_lengths, num_exprs, (int64_t)T, c.nv, prog_off, progs, prog_len, size_s,
                           need_s, frames, (int32_t)T, status);
        he = hipGetLastError();
    }
    if (he == hipSuccess) he = launch_order_by_len(prog_len, num_exprs, order, s);
    if (he != hipSuccess) return c.hip_fail(he, "sr_eval_mse_noalloc compile");
    BatchArgs a{};
    a.progs = progs;
    a.prog_off = prog_off;
    a.prog_len = prog_len;
    a.order = order;
    const EvalBufs bufs{psum, pcnt, counters, d_mse, (long long *)d_valid};
    return enqueue_batched(c, MODE_MSE, lp, a, num_exprs, s, bufs);
}
