// sr_kernels.h -- evaluation kernels (templated; instantiated per output mode
// in sr_eval_*.hip so the variants compile in parallel).
//
// Data layout in HBM:
//   X      fp32 column-major [nv+1][N], y = column nv (cl_gpu_simple.cu:79-84)
//   progs  sr::Ins per expression, packed; prog_off/prog_len per expression
//   cache  fp32 [slots][N] subtree value columns (cl_gpu_common.h:76-83)
//   psum/pcnt  per-(expression, partial slot) fp64 sum of squared error and
//              count of finite predictions; reduced by k_finalize_mse
#pragma once

#include "sr_device.h"

namespace sr {

// MODE_MSE_REF: MSE of skeletons whose OP_REF leaves read cache columns and
// whose STORE_CODE instructions write the accumulator to a column (a.out is
// the column base); the other modes never see either, so their instruction
// loop has no memory access (and the next tile's X loads stay in flight
// across it).
enum Mode { MODE_MSE = 0, MODE_PRED = 1, MODE_COLUMN = 2, MODE_MSE_REF = 3 };
constexpr bool is_mse_mode(int m) { return m == MODE_MSE || m == MODE_MSE_REF; }
constexpr bool is_column_mode(int m) { return m == MODE_COLUMN; }
constexpr bool has_ref(int m) { return m == MODE_MSE_REF; }

constexpr int kBlock = 256;  // 4 waves per workgroup, one per SIMD
constexpr int kWave = 64;

struct BatchArgs {
    const Ins *progs;
    const int32_t *prog_off;  // null: fixed stride (device-resident GA population)
    const int32_t *prog_len;
    const int32_t *order;     // expression id for each rank (longest first)
    const int32_t *out_col;   // MODE_COLUMN: destination cache slot per expression
    const float *X;
    const float *cache;       // OP_REF source columns (may be null)
    const int32_t *ref_pf;    // MODE_MSE_REF: [E][kRefPrefetch] prefetched columns, -1 = none
    float *out;               // MODE_PRED: [E][N]; MODE_COLUMN: cache base
    double *psum;             // MODE_MSE: [E][P]
    long long *pcnt;
    uint32_t *counters;       // one work counter per XCD queue
    int64_t N;
    int64_t slice;            // datapoints per XCD queue (multiple of 64)
    int64_t chunk;            // datapoints per work item
    int n_exprs;
    const int *n_exprs_dev;   // when set: the program count, read on the device (n_exprs = capacity)
    int nv;
    int n_queues;             // 1 or 8 (one per XCD)
    int chunks_per_queue;
    int P;                    // partial slots per expression = n_queues*chunks_per_queue
    int prog_stride;          // programs at e*prog_stride when prog_off is null
    float *spill;             // vector interpreter: per-wave stack spill area (k_vec)
    // MODE_MSE_REF with k_vec (may be null = every program): need_mse[e] == 0
    // marks a program that only writes columns; it loads no target, computes
    // no squared errors and leaves zero partials (MSE NaN, count 0).
    const int32_t *need_mse;
};

// ---------------------------------------------------------------------------
// Batched persistent interpreter: each wave pulls (expression, datapoint chunk)
// work items from its XCD's queue; items run chunk by chunk, longest
// expression first within a chunk.
// Lane l of a wave evaluates datapoints tile + d*64 + l for d < DPT, so every
// X / y / cache access is a coalesced 256-byte wave load.
// ---------------------------------------------------------------------------
// One instruction of the accumulator program on every datapoint a lane owns.
// cv is the value of a constant operand (the literal's bits, or a free
// constant of a parametric program); ref(l...) loads OP_REF operands into
// l.o (only reached when REF).  Stack traffic sits in top-level ifs of its
// own (a stack write nested deeper in the decode makes the compiler copy whole
// stack arrays between branches); constants and variables are fetched
// branch-free.
template <bool REF, typename RefLoad, typename... L>
SR_DEV void interp_step(int code, int arg, float cv, int &sp, RefLoad &&ref, L &...l)
{
    if (code & BC_PUSH) {
        ((l.stk[sp] = l.top), ...);
        ++sp;
    }
    const int src = bc_src(code);
    if (src == SRC_POP) {
        --sp;
        ((l.o = l.stk[sp]), ...);
    } else if (REF && src == SRC_REF) {
        ref(l...);
    } else {
        const bool is_var = src == SRC_VAR;
        const int xi = is_var ? arg : 0;
        for_each_lane([&](int, auto &ll) { ll.o = is_var ? ll.x[xi] : cv; }, l...);
    }
    const int cls = bc_cls(code);
    if (cls != CL_BIN) {
        if (src != SRC_NONE) ((l.top = l.o), ...);
        if (cls == CL_UN) apply_unary(bc_op(code), l...);
    } else {
        if (code & BC_SWAP) {
            for_each_lane(
                [&](int, auto &ll) {
                    float t = ll.top;
                    ll.top = ll.o;
                    ll.o = t;
                },
                l...);
        }
        apply_binary(bc_op(code), l...);
    }
}

// Program words staged per wave in LDS: the longest program plus one word
// read ahead past its end.
constexpr int kProgLds = MAX_EXPR_LEN + 1;

// Waves per SIMD the register allocation must leave room for.  Occupancy
// decides this interpreter: the 3-datapoint variant runs 4 waves per SIMD
// (<= 128 VGPRs; the cache-skeleton mode spills a few) in every mode; the
// 4-datapoint variant fits 3.
constexpr int batched_min_waves(int dpt, int) { return dpt == 3 ? 4 : 1; }

template <int DPT, int S, int XC, int MODE>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(batched_min_waves(DPT, MODE))))
void k_batched(BatchArgs a)
{
    // Each wave copies its expression's program into LDS once per work item
    // (coalesced), then reads instruction i + 1 from LDS while executing
    // instruction i, so dispatch never waits on memory.  X (and y) of the next
    // datapoint tile are loaded while the current tile runs.
    __shared__ Ins s_prog[kBlock / kWave][kProgLds];
    const int lane = threadIdx.x & (kWave - 1);
    Ins *lp = s_prog[threadIdx.x / kWave];
    const int q = a.n_queues > 1 ? (int)(blockIdx.x % (unsigned)a.n_queues) : 0;
    const int n_exprs = a.n_exprs_dev ? min(*a.n_exprs_dev, a.n_exprs) : a.n_exprs;
    const int items = n_exprs * a.chunks_per_queue;
    const int64_t N = a.N;
    constexpr int64_t kTile = (int64_t)kWave * DPT;
    // OP_REF operands held in registers (current and next tile): 2 * KR * DPT
    // VGPRs, so fewer at higher DPT.
    constexpr int KR = has_ref(MODE) ? (DPT >= 3 ? 2 : 4) : 0;
    float rr[KR > 0 ? KR : 1][DPT], rn[KR > 0 ? KR : 1][DPT];

    with_lanes<DPT, S, XC>([&](auto &...L) {
        for (;;) {
            int item = 0;
            if (lane == 0) item = (int)atomicAdd(&a.counters[q], 1u);
            item = __builtin_amdgcn_readfirstlane(item);
            if (item >= items) break;
            // chunk-major: the waves of a queue work on one datapoint chunk
            // at a time, so its X columns stay in the XCD's L2 across
            // expressions; within a chunk, longest expression first.
            const int chunk = item / n_exprs;
            const int rank = item - chunk * n_exprs;
            const int e = a.order[rank];
            const Ins *prog =
                a.progs + (a.prog_off ? (int64_t)a.prog_off[e] : (int64_t)e * a.prog_stride);
            const int n = a.prog_len[e];
            for (int k = lane; k < n; k += kWave) lp[k] = prog[k];
            // columns of the first KR OP_REF operands, loaded a tile ahead
            int pf[KR > 0 ? KR : 1];
#pragma unroll
            for (int j = 0; j < KR; ++j) pf[j] = __builtin_amdgcn_readfirstlane(a.ref_pf[(int64_t)e * kRefPrefetch + j]);
            __builtin_amdgcn_wave_barrier();
            int64_t q0 = (int64_t)q * a.slice;
            int64_t q1 = q0 + a.slice < N ? q0 + a.slice : N;
            int64_t c0 = q0 + (int64_t)chunk * a.chunk;
            int64_t c1 = c0 + a.chunk < q1 ? c0 + a.chunk : q1;
            double sum = 0.0;
            long long cnt = 0;
            float *col_out = nullptr;
            if constexpr (is_column_mode(MODE)) col_out = a.out + (int64_t)a.out_col[e] * N;
            if constexpr (MODE == MODE_PRED) col_out = a.out + (int64_t)e * N;

            // Tile addressing: a uniform 64-bit base per column and tile plus
            // a 32-bit lane offset (clamped to the tile's last valid
            // datapoint), so every access is one scaled-offset load.
            auto tile_last = [&](int64_t t) { return (unsigned)((c1 - t < kTile ? c1 - t : kTile) - 1); };
            auto load_tile = [&](int64_t t) {
                const unsigned last = tile_last(t);
                for_each_lane(
                    [&](int d, auto &l) {
                        const unsigned i = min((unsigned)(d * kWave + lane), last);
#pragma unroll
                        for (int k = 0; k < XC; ++k)
                            if (k < a.nv) l.xn[k] = (a.X + (int64_t)k * N + t)[i];
                        if constexpr (is_mse_mode(MODE)) l.yn = (a.X + (int64_t)a.nv * N + t)[i];
#pragma unroll
                        for (int j = 0; j < KR; ++j)
                            if (pf[j] >= 0) rn[j][d] = (a.cache + (int64_t)pf[j] * N + t)[i];
                    },
                    L...);
            };
            if (c0 < c1) load_tile(c0);
            for (int64_t t0 = c0; t0 < c1; t0 += kTile) {
                for_each_lane(
                    [&](int d, auto &l) {
#pragma unroll
                        for (int k = 0; k < XC; ++k) l.x[k] = l.xn[k];
                        l.y = l.yn;
#pragma unroll
                        for (int j = 0; j < KR; ++j) rr[j][d] = rn[j][d];
                    },
                    L...);
                if (t0 + kTile < c1) load_tile(t0 + kTile);
                const unsigned last = tile_last(t0);
                int sp = 0;
                int code = __builtin_amdgcn_readfirstlane(lp[0].code);
                int arg = __builtin_amdgcn_readfirstlane(lp[0].arg);
                auto ref = [&](auto &...ll) {
                    const int j = bc_pf(code) - 1;  // uniform
                    if (j >= 0 && j < KR) {
                        for_each_lane(
                            [&](int d, auto &l) {
                                float v = rr[0][d];
#pragma unroll
                                for (int k = 1; k < KR; ++k)
                                    if (j == k) v = rr[k][d];
                                l.o = v;
                            },
                            ll...);
                    } else {
                        const float *colr = a.cache + (int64_t)arg * N + t0;
                        for_each_lane(
                            [&](int d, auto &l) { l.o = colr[min((unsigned)(d * kWave + lane), last)]; },
                            ll...);
                    }
                };
                for (int i = 0; i < n; ++i) {
                    const Ins nx = lp[i + 1];
                    // OP_STORE runs through the step as a unary no-op (no
                    // operand, top kept), then writes top to its column: a
                    // branch around the whole step would make the compiler
                    // copy the stack registers between the two paths.
                    interp_step<has_ref(MODE)>(code, arg, __int_as_float(arg), sp, ref, L...);
                    if (has_ref(MODE) && code == STORE_CODE) {
                        float *cs = a.out + (int64_t)arg * N + t0;
                        for_each_lane(
                            [&](int d, auto &l) {
                                const unsigned i = (unsigned)(d * kWave + lane);
                                if (i <= last) cs[i] = l.top;
                            },
                            L...);
                    }
                    code = __builtin_amdgcn_readfirstlane(nx.code);
                    arg = __builtin_amdgcn_readfirstlane(nx.arg);
                }
                // Epilogue.
                for_each_lane(
                    [&](int d, auto &l) {
                        const unsigned i = (unsigned)(d * kWave + lane);
                        if constexpr (is_mse_mode(MODE)) {
                            bool ok = i <= last && isfinite(l.top);
                            float diff = l.top - l.y;
                            float sq = diff * diff;
                            sum += ok ? (double)sq : 0.0;
                            cnt += ok ? 1 : 0;
                        } else {
                            if (i <= last) (col_out + t0)[i] = l.top;
                        }
                    },
                    L...);
            }
            if constexpr (is_mse_mode(MODE)) {
                sum = wave_sum(sum);
                cnt = wave_sum(cnt);
                if (lane == 0) {
                    int slot = q * a.chunks_per_queue + chunk;
                    a.psum[(int64_t)e * a.P + slot] = sum;
                    a.pcnt[(int64_t)e * a.P + slot] = cnt;
                }
            }
            // the next item restages lp: finish this item's reads first
            __builtin_amdgcn_wave_barrier();
        }
    });
}

// ---------------------------------------------------------------------------
// Plain interpreter: one launch per expression, one datapoint per thread, raw
// prefix tokens read right to left (cl_gpu_simple.cu:86-119 restated), MSE
// fused: one (sum, count) partial per workgroup.
// ---------------------------------------------------------------------------
template <int S, int XC>
__global__ __launch_bounds__(kBlock) void k_plain(const Ins *prog, int len, const float *X,
                                                  int64_t N, int nv, double *psum,
                                                  long long *pcnt)
{
    const int64_t dp = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t dc = dp < N ? dp : N - 1;
    double sum = 0.0;
    long long cnt = 0;
    with_lanes<1, S, XC>([&](auto &l) {
#pragma unroll
        for (int k = 0; k < XC; ++k)
            if (k < nv) l.x[k] = X[(int64_t)k * N + dc];
        l.top = 0.0f;
        int sp = 0;
        for (int i = len - 1; i >= 0; --i) {
            const Ins ins = prog[i];
            const int t = ins.code;
            if (t <= 0) {
                l.stk[sp] = l.top;  // the first leaf pushes an unused value
                ++sp;
                if (t == TOK_CONST) l.top = __int_as_float(ins.arg);
                else l.top = l.x[ins.arg];
            } else if (t < 10) {
                --sp;
                l.o = l.stk[sp];
                apply_binary(t, l);
            } else {
                apply_unary(t, l);
            }
        }
        float y = X[(int64_t)nv * N + dc];
        bool ok = dp < N && isfinite(l.top);
        float diff = l.top - y;
        float sq = diff * diff;
        sum = ok ? (double)sq : 0.0;
        cnt = ok ? 1 : 0;
    });
    __shared__ double s_sum[kBlock / kWave];
    __shared__ long long s_cnt[kBlock / kWave];
    sum = wave_sum(sum);
    cnt = wave_sum(cnt);
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        s_sum[w] = sum;
        s_cnt[w] = cnt;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        long long c = 0;
        for (int k = 0; k < kBlock / kWave; ++k) {
            t += s_sum[k];
            c += s_cnt[k];
        }
        psum[blockIdx.x] = t;
        pcnt[blockIdx.x] = c;
    }
}

}  // namespace sr
