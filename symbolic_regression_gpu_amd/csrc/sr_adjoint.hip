// sr_adjoint.hip -- parametric forward (with finite-difference replicas) and
// reverse-mode adjoint kernels for constant fitting.
#include "sr_adjoint.h"
#include "sr_adjoint_prog.h"
#include "sr_launch.h"
#include "sr_param_launch.h"

namespace sr {

// Copy expression e's program into this wave's LDS slot, with every free
// constant's current value (plus the finite-difference step on slot pert)
// written into its operand word, so the interpreter reads it as a literal.
SR_DEV void stage_param_program(const ParamArgs &a, int e, int n, int lane, int pert, float delta, Ins *lp,
                                int32_t *la)
{
    const Ins *prog = a.progs + a.prog_off[e];
    const int32_t *aux = a.aux ? a.aux + a.prog_off[e] : nullptr;
    for (int k = lane; k < n; k += kWave) {
        Ins ins = prog[k];
        const int s = bc_src(ins.code) == SRC_CONST ? bc_slot(ins.code) : -1;
        if (s >= 0) {
            const float c = a.cvals[(int64_t)e * MAX_SLOTS + s];
            ins.arg = __float_as_int(s == pert ? c + delta : c);
        }
        lp[k] = ins;
        if (la) la[k] = aux[k];
    }
    __builtin_amdgcn_wave_barrier();
}

// ---------------------------------------------------------------------------
// Forward: MSE of every (replica, expression).  Replica r of a
// central-difference batch perturbs free constant r/2 by +h (r even) or -h.
// Same interpreter as the batched evaluation (program in LDS, one-ahead
// instruction reads, next tile's X loaded during the current one).
// ---------------------------------------------------------------------------
template <int DPT, int XC>
__global__ __launch_bounds__(kBlock) void k_param_fwd(ParamArgs a)
{
    extern __shared__ Ins s_prog[];  // [waves][prog_lds]
    const int lane = threadIdx.x & (kWave - 1);
    Ins *lp = s_prog + (threadIdx.x / kWave) * a.prog_lds;
    const int q = a.n_queues > 1 ? (int)(blockIdx.x % (unsigned)a.n_queues) : 0;
    const int per_rep = a.n_exprs * a.chunks_per_queue;
    const int items = a.replicas * per_rep;
    const int64_t N = a.N;
    constexpr int64_t kTile = (int64_t)kWave * DPT;
    with_lanes<DPT, kStackRegs, XC>([&](auto &...L) {
        for (;;) {
            int item = 0;
            if (lane == 0) item = (int)atomicAdd(&a.counters[q], 1u);
            item = __builtin_amdgcn_readfirstlane(item);
            if (item >= items) break;
            const int rep = item / per_rep;
            const int rem = item - rep * per_rep;
            const int chunk = rem / a.n_exprs;  // chunk-major (X stays in L2)
            const int rank = rem - chunk * a.n_exprs;
            const int e = a.order[rank];
            const int pert = a.replicas > 1 ? rep >> 1 : -1;
            const float delta = a.replicas > 1 ? ((rep & 1) ? -a.fd_h : a.fd_h) : 0.0f;
            const bool skip = pert >= a.nslots[e];
            const int n = a.prog_len[e];
            const int64_t q0 = (int64_t)q * a.slice;
            const int64_t q1 = q0 + a.slice < N ? q0 + a.slice : N;
            const int64_t c0 = q0 + (int64_t)chunk * a.chunk;
            const int64_t c1 = skip ? c0 : (c0 + a.chunk < q1 ? c0 + a.chunk : q1);
            double sum = 0.0;
            long long cnt = 0;
            if (!skip) stage_param_program(a, e, n, lane, pert, delta, lp, nullptr);
            auto load_tile = [&](int64_t t) {
                for_each_lane(
                    [&](int d, auto &l) {
                        int64_t dp = t + d * kWave + lane;
                        int64_t dc = dp < N ? dp : N - 1;
#pragma unroll
                        for (int k = 0; k < XC; ++k)
                            if (k < a.nv) l.xn[k] = a.X[(int64_t)k * N + dc];
                        l.yn = a.X[(int64_t)a.nv * N + dc];
                    },
                    L...);
            };
            auto noref = [](auto &...) {};
            if (c0 < c1) load_tile(c0);
            for (int64_t t0 = c0; t0 < c1; t0 += kTile) {
                for_each_lane(
                    [&](int, auto &l) {
#pragma unroll
                        for (int k = 0; k < XC; ++k) l.x[k] = l.xn[k];
                        l.y = l.yn;
                    },
                    L...);
                if (t0 + kTile < c1) load_tile(t0 + kTile);
                int sp = 0;
                int code = __builtin_amdgcn_readfirstlane(lp[0].code);
                int arg = __builtin_amdgcn_readfirstlane(lp[0].arg);
                for (int i = 0; i < n; ++i) {
                    const Ins nx = lp[i + 1];
                    interp_step<false>(code, arg, __int_as_float(arg), sp, noref, L...);
                    code = __builtin_amdgcn_readfirstlane(nx.code);
                    arg = __builtin_amdgcn_readfirstlane(nx.arg);
                }
                for_each_lane(
                    [&](int d, auto &l) {
                        int64_t dp = t0 + d * kWave + lane;
                        bool ok = dp < c1 && isfinite(l.top);
                        float diff = l.top - l.y;
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
            __builtin_amdgcn_wave_barrier();
        }
    });
}

// ---------------------------------------------------------------------------
// Adjoint.  Per tile of 64 * DPT datapoints: the forward pass stores every
// value the reverse pass reads on this wave's tape (rows assigned on the
// host; one 64 * DPT float row each, lane-major so a lane's DPT values are
// one vector store); the reverse pass walks the active instructions from the
// last one down, jumping over subtrees without a free constant, with an
// adjoint stack mirroring the value stack (a popped operand's adjoint is
// pushed and returned at the instruction that pushed the value).
// ---------------------------------------------------------------------------
constexpr int kAdjRegs = 9;  // > 8 floats: uniform dynamic indexing uses GPR-index mode

template <int XC>
struct AdjLane {
    float (&stk)[kStackRegs];
    float (&x)[kXRegs<XC>];
    float &top;
    float &o;
    float (&astk)[kAdjRegs];
    float (&tg)[kAdjRegs];  // per-slot gradient sums of the work item (masked per datapoint)
    float &at;
};

template <int DPT, int XC, typename F, typename... L>
SR_DEV void with_adj_lanes(F &&f, L &...ls)
{
    if constexpr (sizeof...(L) == DPT) {
        f(ls...);
    } else {
        float stk[kStackRegs];
        float x[kXRegs<XC>];
        float top = 0.0f, o = 0.0f, at = 0.0f;
        float astk[kAdjRegs], tg[kAdjRegs];
        AdjLane<XC> l{stk, x, top, o, astk, tg, at};
        with_adj_lanes<DPT, XC>(f, ls..., l);
    }
}

// Tape row r of this wave: 64 lanes x DPT floats, lane-major.
template <int DPT>
struct WaveTape {
    float *base;  // this wave's region
    int lane;
    template <typename... L>
    SR_DEV void store(int r, L &...l) const
    {
        float *p = base + ((int64_t)r * kWave + lane) * DPT;
        if constexpr (DPT == 4) {
            float4 v;
            float t[4];
            int d = 0;
            ((t[d++] = l.top), ...);
            v.x = t[0];
            v.y = t[1];
            v.z = t[2];
            v.w = t[3];
            *reinterpret_cast<float4 *>(p) = v;
        } else if constexpr (DPT == 2) {
            float t[2];
            int d = 0;
            ((t[d++] = l.top), ...);
            *reinterpret_cast<float2 *>(p) = make_float2(t[0], t[1]);
        } else {
            ((*p = l.top), ...);
        }
    }
    SR_DEV void load(int r, float (&v)[DPT]) const
    {
        const float *p = base + ((int64_t)r * kWave + lane) * DPT;
        if constexpr (DPT == 4) {
            const float4 t = *reinterpret_cast<const float4 *>(p);
            v[0] = t.x;
            v[1] = t.y;
            v[2] = t.z;
            v[3] = t.w;
        } else if constexpr (DPT == 2) {
            const float2 t = *reinterpret_cast<const float2 *>(p);
            v[0] = t.x;
            v[1] = t.y;
        } else {
            v[0] = *p;
        }
    }
};

// At least 4 waves per SIMD (<= 128 VGPRs): the 2-datapoint variants
// otherwise take 172-181 VGPRs (2 waves); the kernel is latency-bound, and a
// few spilled registers cost less than the lost occupancy (measured on 4096
// expressions x 256k: 253 ms per gradient at 2 waves, 183 at 3, 153 at 4).
template <int DPT, int XC>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_adjoint(ParamArgs a)
{
    extern __shared__ Ins s_prog[];  // [waves][prog_lds] programs, then [waves][prog_lds] aux words
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    Ins *lp = s_prog + w * a.prog_lds;
    int32_t *la = reinterpret_cast<int32_t *>(s_prog + (kBlock / kWave) * a.prog_lds) + w * a.prog_lds;
    const int q = a.n_queues > 1 ? (int)(blockIdx.x % (unsigned)a.n_queues) : 0;
    const int64_t wave = (int64_t)blockIdx.x * (kBlock / kWave) + w;
    const WaveTape<DPT> tape{a.tape + wave * (int64_t)a.tape_stride * kWave * DPT, lane};
    const int items = a.n_exprs * a.chunks_per_queue;
    const int64_t N = a.N;
    constexpr int64_t kTile = (int64_t)kWave * DPT;

    with_adj_lanes<DPT, XC>([&](auto &...L) {
        for (;;) {
            int item = 0;
            if (lane == 0) item = (int)atomicAdd(&a.counters[q], 1u);
            item = __builtin_amdgcn_readfirstlane(item);
            if (item >= items) break;
            const int chunk = item / a.n_exprs;  // chunk-major (X stays in L2)
            const int rank = item - chunk * a.n_exprs;
            const int e = a.order[rank];
            const int n = a.prog_len[e];
            stage_param_program(a, e, n, lane, -1, 0.0f, lp, la);
            const int64_t q0 = (int64_t)q * a.slice;
            const int64_t q1 = q0 + a.slice < N ? q0 + a.slice : N;
            const int64_t c0 = q0 + (int64_t)chunk * a.chunk;
            const int64_t c1 = c0 + a.chunk < q1 ? c0 + a.chunk : q1;
            double sum = 0.0;
            long long cnt = 0;
            for_each_lane(
                [&](int, auto &l) {
#pragma unroll
                    for (int k = 0; k < kAdjRegs; ++k) l.tg[k] = 0.0f;
                },
                L...);
            auto noref = [](auto &...) {};
            for (int64_t t0 = c0; t0 < c1; t0 += kTile) {
                float yv[DPT];
                bool okv[DPT];
                for_each_lane(
                    [&](int d, auto &l) {
                        int64_t dp = t0 + d * kWave + lane;
                        int64_t dc = dp < N ? dp : N - 1;
#pragma unroll
                        for (int k = 0; k < XC; ++k)
                            if (k < a.nv) l.x[k] = a.X[(int64_t)k * N + dc];
                        yv[d] = a.X[(int64_t)a.nv * N + dc];
                    },
                    L...);
                // forward, taping the values the reverse pass reads
                int sp = 0;
                int code = __builtin_amdgcn_readfirstlane(lp[0].code);
                int arg = __builtin_amdgcn_readfirstlane(lp[0].arg);
                for (int i = 0; i < n; ++i) {
                    const Ins nx = lp[i + 1];
                    interp_step<false>(code, arg, __int_as_float(arg), sp, noref, L...);
                    const int tr = bc_tape(code);
                    if (tr >= 0) tape.store(tr, L...);
                    code = __builtin_amdgcn_readfirstlane(nx.code);
                    arg = __builtin_amdgcn_readfirstlane(nx.arg);
                }
                bool any = false;
                for_each_lane(
                    [&](int d, auto &l) {
                        int64_t dp = t0 + d * kWave + lane;
                        okv[d] = dp < c1 && isfinite(l.top);
                        float diff = l.top - yv[d];
                        float sq = diff * diff;
                        sum += okv[d] ? (double)sq : 0.0;
                        cnt += okv[d] ? 1 : 0;
                        l.at = okv[d] ? diff : 0.0f;
                        any = any || okv[d];
                    },
                    L...);
                if (!__any(any)) continue;
                // reverse over the active instructions
                int asp = 0;
                int i = n - 1;
                while (i >= 0) {
                    const int rc = __builtin_amdgcn_readfirstlane(lp[i].code);
                    const int ra = __builtin_amdgcn_readfirstlane(lp[i].arg);
                    const int rx = __builtin_amdgcn_readfirstlane(la[i]);
                    if (!(rc & BC_ACTIVE)) {
                        // a subtree without free constants: skip [rx, i]
                        const int sc = __builtin_amdgcn_readfirstlane(lp[rx].code);
                        if (sc & BC_PUSH) {
                            --asp;
                            ((L.at = L.astk[asp]), ...);
                        } else {
                            ((L.at = 0.0f), ...);
                        }
                        i = rx - 1;
                        continue;
                    }
                    const int src = bc_src(rc);
                    const int op = bc_op(rc);
                    float yo[DPT], xi[DPT], pv[DPT];
                    tape.load(bc_tape(rc), yo);
                    if (rx >= 0) tape.load(rx, xi);
                    if (src == SRC_POP) tape.load(ra, pv);
                    const float cv = __int_as_float(ra);
                    const int slot = src == SRC_CONST ? bc_slot(rc) : -1;
                    for_each_lane(
                        [&](int d, auto &l) {
                            const float x_in = rx >= 0 ? xi[d] : 0.0f;
                            float ov = src == SRC_POP ? pv[d] : (src == SRC_VAR ? l.x[ra] : cv);
                            float a_o = 0.0f, a_in = 0.0f;
                            const int cls = bc_cls(rc);
                            if (cls == CL_LOAD) {
                                a_o = l.at;
                            } else if (cls == CL_BIN) {
                                const bool sw = (rc & BC_SWAP) != 0;
                                const float A = sw ? ov : x_in;
                                const float B = sw ? x_in : ov;
                                float dA, dB;
                                d_binary(op, A, B, yo[d], dA, dB);
                                const float gA = l.at * dA, gB = l.at * dB;
                                a_o = sw ? gA : gB;
                                a_in = sw ? gB : gA;
                            } else {
                                const float u = src != SRC_NONE ? ov : x_in;
                                const float g = l.at * d_unary(op, u, yo[d]);
                                if (src != SRC_NONE) a_o = g;
                                else a_in = g;
                            }
                            l.o = a_o;
                            l.at = a_in;
                        },
                        L...);
                    if (slot >= 0) {
                        // a point outside the MSE adds nothing, not 0 * inf
                        for_each_lane([&](int d, auto &l) { l.tg[slot] += okv[d] ? l.o : 0.0f; }, L...);
                    } else if (src == SRC_POP) {
                        ((L.astk[asp] = L.o), ...);
                        ++asp;
                    }
                    if (rc & BC_PUSH) {
                        --asp;
                        ((L.at = L.astk[asp]), ...);
                    }
                    --i;
                }
            }
            sum = wave_sum(sum);
            cnt = wave_sum(cnt);
            double g[MAX_SLOTS];
#pragma unroll
            for (int k = 0; k < MAX_SLOTS; ++k) {
                double t = 0.0;
                for_each_lane([&](int, auto &l) { t += (double)l.tg[k]; }, L...);
                g[k] = wave_sum(t);
            }
            if (lane == 0) {
                const int slot = q * a.chunks_per_queue + chunk;
                a.psum[(int64_t)e * a.P + slot] = sum;
                a.pcnt[(int64_t)e * a.P + slot] = cnt;
#pragma unroll
                for (int k = 0; k < MAX_SLOTS; ++k) a.pgrad[((int64_t)e * a.P + slot) * MAX_SLOTS + k] = g[k];
            }
            __builtin_amdgcn_wave_barrier();
        }
    });
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
        hipLaunchKernelGGL((k_param_fwd<D, XCV>), dim3(grid), dim3(kBlock), param_lds_bytes(a.prog_lds, false), s, a); \
        return hipGetLastError();                                                                \
    }
    SR_PF(1, 4) SR_PF(2, 4) SR_PF(4, 4) SR_PF(1, 8) SR_PF(2, 8) SR_PF(4, 8) SR_PF(1, 16) SR_PF(2, 16)
    SR_PF(4, 16) SR_PF(1, 32)
#undef SR_PF
    return hipErrorInvalidValue;
}

int param_fwd_occupancy(int dpt, int xc, int prog_lds)
{
    int b = 0;
#define SR_PF(D, XCV)                                                                          \
    if (dpt == D && xc == XCV)                                                                 \
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, (const void *)k_param_fwd<D, XCV>, \
                                                            kBlock, param_lds_bytes(prog_lds, false)) == hipSuccess ? b : 0;
    SR_PF(1, 4) SR_PF(2, 4) SR_PF(4, 4) SR_PF(1, 8) SR_PF(2, 8) SR_PF(4, 8) SR_PF(1, 16) SR_PF(2, 16)
    SR_PF(4, 16) SR_PF(1, 32)
#undef SR_PF
    return 0;
}

hipError_t launch_adjoint(int dpt, int xc, const ParamArgs &a, int grid, hipStream_t s)
{
#define SR_AD(D, XCV)                                                                    \
    if (dpt == D && xc == XCV) {                                                         \
        hipLaunchKernelGGL((k_adjoint<D, XCV>), dim3(grid), dim3(kBlock), param_lds_bytes(a.prog_lds, true), s, a); \
        return hipGetLastError();                                                        \
    }
    SR_AD(1, 4) SR_AD(2, 4) SR_AD(1, 8) SR_AD(2, 8) SR_AD(1, 16) SR_AD(2, 16) SR_AD(1, 32)
#undef SR_AD
    return hipErrorInvalidValue;
}

int adjoint_occupancy(int dpt, int xc, int prog_lds)
{
    int b = 0;
#define SR_AD(D, XCV)                                                                             \
    if (dpt == D && xc == XCV)                                                                    \
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, (const void *)k_adjoint<D, XCV>,  \
                                                            kBlock, param_lds_bytes(prog_lds, true)) == hipSuccess ? b : 0;
    SR_AD(1, 4) SR_AD(2, 4) SR_AD(1, 8) SR_AD(2, 8) SR_AD(1, 16) SR_AD(2, 16) SR_AD(1, 32)
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
