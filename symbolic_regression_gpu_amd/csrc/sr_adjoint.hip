// sr_adjoint.hip -- parametric forward (with finite-difference replicas) and
// reverse-mode adjoint kernels for constant fitting.
#include "sr_adjoint.h"
#include "sr_launch.h"
#include "sr_param_launch.h"

namespace sr {

// Constant operand of a parametric program.
SR_DEV float param_const(const ParamArgs &a, int code, int32_t arg, int e, int pert_slot, float delta)
{
    const int s = bc_slot(code);
    if (s < 0) return __int_as_float(arg);
    const float c = a.cvals[(int64_t)e * MAX_SLOTS + s];
    return s == pert_slot ? c + delta : c;
}

// ---------------------------------------------------------------------------
// Forward: MSE of every (replica, expression).  Replica r of a
// central-difference batch perturbs free constant r/2 by +h (r even) or -h.
// ---------------------------------------------------------------------------
template <int DPT, int XC>
__global__ __launch_bounds__(kBlock) void k_param_fwd(ParamArgs a)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int q = a.n_queues > 1 ? (int)(blockIdx.x % (unsigned)a.n_queues) : 0;
    const int per_rep = a.n_exprs * a.chunks_per_queue;
    const int items = a.replicas * per_rep;
    const int64_t N = a.N;
    with_lanes<DPT, kStackSlots, XC>([&](auto &...L) {
        for (;;) {
            int item = 0;
            if (lane == 0) item = (int)atomicAdd(&a.counters[q], 1u);
            item = __shfl(item, 0, kWave);
            if (item >= items) break;
            const int rep = item / per_rep;
            const int rem = item - rep * per_rep;
            const int rank = rem / a.chunks_per_queue;
            const int chunk = rem - rank * a.chunks_per_queue;
            const int e = a.order[rank];
            const int pert = a.replicas > 1 ? rep >> 1 : -1;
            const float delta = a.replicas > 1 ? ((rep & 1) ? -a.fd_h : a.fd_h) : 0.0f;
            const bool skip = pert >= a.nslots[e];
            const Ins *prog = a.progs + a.prog_off[e];
            const int n = a.prog_len[e];
            const int64_t q0 = (int64_t)q * a.slice;
            const int64_t q1 = q0 + a.slice < N ? q0 + a.slice : N;
            const int64_t c0 = q0 + (int64_t)chunk * a.chunk;
            const int64_t c1 = skip ? c0 : (c0 + a.chunk < q1 ? c0 + a.chunk : q1);
            double sum = 0.0;
            long long cnt = 0;
            for (int64_t t0 = c0; t0 < c1; t0 += (int64_t)kWave * DPT) {
                for_each_lane(
                    [&](int d, auto &l) {
                        int64_t dp = t0 + d * kWave + lane;
                        int64_t dc = dp < N ? dp : N - 1;
#pragma unroll
                        for (int k = 0; k < XC; ++k)
                            if (k < a.nv) l.x[k] = a.X[(int64_t)k * N + dc];
                    },
                    L...);
                int sp = 0;
                for (int i = 0; i < n; ++i) {
                    const Ins ins = prog[i];
                    const int code = ins.code;
                    if (code & BC_PUSH) {
                        ((L.stk[sp] = L.top), ...);
                        ++sp;
                    }
                    const int src = bc_src(code);
                    const int op = bc_op(code);
                    if (src == SRC_CONST) {
                        const float c = param_const(a, code, ins.arg, e, pert, delta);
                        ((L.o = c), ...);
                    } else if (src == SRC_VAR) {
                        const int k = ins.arg;
                        ((L.o = L.x[k]), ...);
                    } else if (src == SRC_POP) {
                        --sp;
                        ((L.o = L.stk[sp]), ...);
                    }
                    if (op == 0) {
                        ((L.top = L.o), ...);
                    } else if (op < 10) {
                        if (code & BC_SWAP)
                            for_each_lane(
                                [&](int, auto &l) {
                                    float t = l.top;
                                    l.top = l.o;
                                    l.o = t;
                                },
                                L...);
                        apply_binary(op, L...);
                    } else {
                        if (src != SRC_NONE) ((L.top = L.o), ...);
                        apply_unary(op, L...);
                    }
                }
                for_each_lane(
                    [&](int d, auto &l) {
                        int64_t dp = t0 + d * kWave + lane;
                        int64_t dc = dp < N ? dp : N - 1;
                        float y = a.X[(int64_t)a.nv * N + dc];
                        bool ok = dp < c1 && isfinite(l.top);
                        float diff = l.top - y;
                        float sq = diff * diff;
                        sum += ok ? (double)sq : 0.0;
                        cnt += ok ? 1 : 0;
                    },
                    L...);
            }
            sum = wave_sum(sum);
            cnt = wave_sum(cnt);
            if (lane == 0) {
                const int64_t row = (int64_t)rep * a.n_exprs + e;
                const int slot = q * a.chunks_per_queue + chunk;
                a.psum[row * a.P + slot] = sum;
                a.pcnt[row * a.P + slot] = cnt;
            }
        }
    });
}

// ---------------------------------------------------------------------------
// Tape: one 256-byte row (64 lanes) per instruction in this wave's region of
// a global scratch buffer; forward stores are fire-and-forget and the reverse
// sweep loads each row one step ahead of its use.
// ---------------------------------------------------------------------------
struct GlobalTape {
    float *base;  // this wave's region, lane-offset; row i at base[i*64]
    SR_DEV float get(int i) const { return base[(int64_t)i * kWave]; }
    SR_DEV void set(int i, float v) { base[(int64_t)i * kWave] = v; }
};

// Accumulator view for apply_unary / apply_binary on a single datapoint.
struct TopO {
    float &top;
    float &o;
};

template <int XC, typename Tape>
SR_DEV void adjoint_items(const ParamArgs &a, Tape &tape, int lane, int q)
{
    const int items = a.n_exprs * a.chunks_per_queue;
    const int64_t N = a.N;
    float stk[kStackSlots], astk[kStackSlots], x[XC], tg[MAX_SLOTS], gacc[MAX_SLOTS];
    for (;;) {
        int item = 0;
        if (lane == 0) item = (int)atomicAdd(&a.counters[q], 1u);
        item = __shfl(item, 0, kWave);
        if (item >= items) break;
        const int rank = item / a.chunks_per_queue;
        const int chunk = item - rank * a.chunks_per_queue;
        const int e = a.order[rank];
        const Ins *prog = a.progs + a.prog_off[e];
        const int n = a.prog_len[e];
        const int64_t q0 = (int64_t)q * a.slice;
        const int64_t q1 = q0 + a.slice < N ? q0 + a.slice : N;
        const int64_t c0 = q0 + (int64_t)chunk * a.chunk;
        const int64_t c1 = c0 + a.chunk < q1 ? c0 + a.chunk : q1;
        double sum = 0.0;
        long long cnt = 0;
#pragma unroll
        for (int k = 0; k < MAX_SLOTS; ++k) gacc[k] = 0.0f;
        for (int64_t t0 = c0; t0 < c1; t0 += kWave) {
            const int64_t dp = t0 + lane;
            const int64_t dc = dp < N ? dp : N - 1;
#pragma unroll
            for (int k = 0; k < XC; ++k)
                if (k < a.nv) x[k] = a.X[(int64_t)k * N + dc];
            // forward with tape
            float top = 0.0f, o = 0.0f;
            int sp = 0;
            for (int i = 0; i < n; ++i) {
                const Ins ins = prog[i];
                const int code = ins.code;
                if (code & BC_PUSH) stk[sp++] = top;
                const int src = bc_src(code);
                const int op = bc_op(code);
                if (src == SRC_CONST) o = param_const(a, code, ins.arg, e, -1, 0.0f);
                else if (src == SRC_VAR) o = x[ins.arg];
                else if (src == SRC_POP) o = stk[--sp];
                if (op == 0) {
                    top = o;
                } else if (op < 10) {
                    if (code & BC_SWAP) {
                        float t = top;
                        top = o;
                        o = t;
                    }
                    TopO l{top, o};
                    apply_binary(op, l);
                } else {
                    if (src != SRC_NONE) top = o;
                    TopO l{top, o};
                    apply_unary(op, l);
                }
                tape.set(i, top);
            }
            const float y = a.X[(int64_t)a.nv * N + dc];
            const bool ok = dp < c1 && isfinite(top);
            const float diff = top - y;
            const float sq = diff * diff;
            sum += ok ? (double)sq : 0.0;
            cnt += ok ? 1 : 0;
            // reverse sweep
            if (__ballot(ok)) {
#pragma unroll
                for (int k = 0; k < MAX_SLOTS; ++k) tg[k] = 0.0f;
                float at = ok ? diff : 0.0f;
                float out_i = top;
                int asp = 0;
                float x_next = n > 1 ? tape.get(n - 2) : 0.0f;
                for (int i = n - 1; i >= 0; --i) {
                    const Ins ins = prog[i];
                    const int code = ins.code;
                    const int src = bc_src(code);
                    const int op = bc_op(code);
                    const float x_in = x_next;  // top before instruction i (tape row i-1)
                    x_next = i > 1 ? tape.get(i - 2) : 0.0f;
                    float ov = 0.0f;
                    if (src == SRC_CONST) ov = param_const(a, code, ins.arg, e, -1, 0.0f);
                    else if (src == SRC_VAR) ov = x[ins.arg];
                    else if (src == SRC_POP) ov = tape.get(ins.arg);
                    float a_o = 0.0f, a_in = 0.0f;
                    if (op == 0) {
                        a_o = at;
                    } else if (op < 10) {
                        const bool sw = (code & BC_SWAP) != 0;
                        const float A = sw ? ov : x_in;
                        const float B = sw ? x_in : ov;
                        float dA, dB;
                        d_binary(op, A, B, out_i, dA, dB);
                        const float gA = at * dA, gB = at * dB;
                        a_o = sw ? gA : gB;
                        a_in = sw ? gB : gA;
                    } else {
                        const float u = src != SRC_NONE ? ov : x_in;
                        const float g = at * d_unary(op, u, out_i);
                        if (src != SRC_NONE) a_o = g;
                        else a_in = g;
                    }
                    if (src == SRC_CONST) {
                        const int s = bc_slot(code);
                        if (s >= 0) tg[s] += a_o;
                    } else if (src == SRC_POP) {
                        astk[asp++] = a_o;
                    }
                    if (code & BC_PUSH) a_in = astk[--asp];
                    at = a_in;
                    out_i = x_in;
                }
#pragma unroll
                for (int k = 0; k < MAX_SLOTS; ++k) gacc[k] += ok ? tg[k] : 0.0f;
            }
        }
        sum = wave_sum(sum);
        cnt = wave_sum(cnt);
        double g[MAX_SLOTS];
#pragma unroll
        for (int k = 0; k < MAX_SLOTS; ++k) g[k] = wave_sum((double)gacc[k]);
        if (lane == 0) {
            const int slot = q * a.chunks_per_queue + chunk;
            a.psum[(int64_t)e * a.P + slot] = sum;
            a.pcnt[(int64_t)e * a.P + slot] = cnt;
#pragma unroll
            for (int k = 0; k < MAX_SLOTS; ++k)
                a.pgrad[((int64_t)e * a.P + slot) * MAX_SLOTS + k] = g[k];
        }
    }
}

template <int XC>
__global__ __launch_bounds__(kBlock) void k_adjoint(ParamArgs a)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int q = a.n_queues > 1 ? (int)(blockIdx.x % (unsigned)a.n_queues) : 0;
    const int64_t wave = (int64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    GlobalTape tape{a.tape + wave * a.tape_stride * kWave + lane};
    adjoint_items<XC>(a, tape, lane, q);
}

// mse[e], and grad[e][k] = 2 * sum_k / count (NaN when no finite point).
__global__ __launch_bounds__(kBlock) void k_finalize_grad(const double *psum, const long long *pcnt,
                                                          const double *pgrad, int P, int E,
                                                          float *mse, float *grad)
{
    const int e = (int)((blockIdx.x * (unsigned)kBlock + threadIdx.x) / kWave);
    const int lane = threadIdx.x & (kWave - 1);
    if (e >= E) return;
    double s = 0.0, g[MAX_SLOTS];
    long long c = 0;
#pragma unroll
    for (int k = 0; k < MAX_SLOTS; ++k) g[k] = 0.0;
    for (int p = lane; p < P; p += kWave) {
        s += psum[(int64_t)e * P + p];
        c += pcnt[(int64_t)e * P + p];
#pragma unroll
        for (int k = 0; k < MAX_SLOTS; ++k) g[k] += pgrad[((int64_t)e * P + p) * MAX_SLOTS + k];
    }
    s = wave_sum(s);
    c = wave_sum(c);
#pragma unroll
    for (int k = 0; k < MAX_SLOTS; ++k) g[k] = wave_sum(g[k]);
    if (lane == 0) {
        mse[e] = c > 0 ? (float)(s / (double)c) : qnan();
#pragma unroll
        for (int k = 0; k < MAX_SLOTS; ++k)
            grad[(int64_t)e * MAX_SLOTS + k] = c > 0 ? (float)(2.0 * g[k] / (double)c) : qnan();
    }
}

hipError_t launch_param_fwd(int dpt, int xc, const ParamArgs &a, int grid, hipStream_t s)
{
#define SR_PF(D, XCV)                                                                            \
    if (dpt == D && xc == XCV) {                                                                \
        hipLaunchKernelGGL((k_param_fwd<D, XCV>), dim3(grid), dim3(kBlock), 0, s, a);            \
        return hipGetLastError();                                                                \
    }
    SR_PF(1, 4) SR_PF(2, 4) SR_PF(4, 4) SR_PF(1, 16) SR_PF(2, 16) SR_PF(4, 16) SR_PF(1, 32)
#undef SR_PF
    return hipErrorInvalidValue;
}

int param_fwd_occupancy(int dpt, int xc)
{
    int b = 0;
#define SR_PF(D, XCV)                                                                          \
    if (dpt == D && xc == XCV)                                                                 \
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, (const void *)k_param_fwd<D, XCV>, \
                                                            kBlock, 0) == hipSuccess ? b : 0;
    SR_PF(1, 4) SR_PF(2, 4) SR_PF(4, 4) SR_PF(1, 16) SR_PF(2, 16) SR_PF(4, 16) SR_PF(1, 32)
#undef SR_PF
    return 0;
}

hipError_t launch_adjoint(int xc, const ParamArgs &a, int grid, hipStream_t s)
{
#define SR_AD(XCV)                                                                    \
    if (xc == XCV) {                                                                  \
        hipLaunchKernelGGL((k_adjoint<XCV>), dim3(grid), dim3(kBlock), 0, s, a);       \
        return hipGetLastError();                                                     \
    }
    SR_AD(4) SR_AD(16) SR_AD(32)
#undef SR_AD
    return hipErrorInvalidValue;
}

int adjoint_occupancy(int xc)
{
    int b = 0;
#define SR_AD(XCV)                                                                              \
    if (xc == XCV)                                                                              \
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, (const void *)k_adjoint<XCV>,   \
                                                            kBlock, 0) == hipSuccess ? b : 0;
    SR_AD(4) SR_AD(16) SR_AD(32)
#undef SR_AD
    return 0;
}

hipError_t launch_finalize_grad(const double *psum, const long long *pcnt, const double *pgrad, int P,
                                int E, float *mse, float *grad, hipStream_t s)
{
    const int per_block = kBlock / kWave;
    hipLaunchKernelGGL(k_finalize_grad, dim3((E + per_block - 1) / per_block), dim3(kBlock), 0, s,
                       psum, pcnt, pgrad, P, E, mse, grad);
    return hipGetLastError();
}

}  // namespace sr

namespace sr {

// Central differences from the replica MSEs: grad[e][k] =
// (mse(c_k + h) - mse(c_k - h)) / ((c_k + h) - (c_k - h)), 0 for unused slots.
__global__ void k_fd_combine(const float *cvals, const float *rep_mse, const int32_t *nslots, int E,
                             float h, float *grad)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= E * MAX_SLOTS) return;
    const int e = t / MAX_SLOTS, k = t - e * MAX_SLOTS;
    if (k >= nslots[e]) {
        grad[t] = 0.0f;
        return;
    }
    const float c = cvals[t];
    const float hi = c + h, lo = c - h;
    const float mp = rep_mse[(int64_t)(2 * k) * E + e];
    const float mm = rep_mse[(int64_t)(2 * k + 1) * E + e];
    grad[t] = (float)(((double)mp - (double)mm) / ((double)hi - (double)lo));
}

hipError_t launch_fd_combine(const float *cvals, const float *rep_mse, const int32_t *nslots, int E,
                             float h, float *grad, hipStream_t s)
{
    const int n = E * MAX_SLOTS;
    hipLaunchKernelGGL(k_fd_combine, dim3((n + 255) / 256), dim3(256), 0, s, cvals, rep_mse, nslots, E,
                       h, grad);
    return hipGetLastError();
}

}  // namespace sr
