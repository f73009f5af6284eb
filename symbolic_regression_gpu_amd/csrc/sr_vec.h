// sr_vec.h -- the vector interpreter: one bytecode instruction at a time over
// a 384-datapoint tile (6 datapoints per lane) per wave.
//
// The lane interpreter (k_batched, sr_kernels.h) holds the variables of its
// tile and the stack in VGPRs, which caps it at 3 datapoints per lane; its
// scalar decode (fetch, source / class / op branches) is then paid once per
// 192 datapoints, and on short cache skeletons the per-tile loads of X|y and
// columns, prefetched one tile ahead, cannot hide their latency.  Here a
// wave decodes an instruction once per 384 datapoints, loads the
// instruction's variable or column operand straight from memory (six
// independent coalesced loads per lane, with the operands of the next two
// instructions in flight while the current one computes), keeps the stack in
// LDS (slots beyond kVecLdsSlots in a per-wave spill area of global memory)
// and fetches instructions with scalar loads.  Nothing depends on the number of
// variables, so one instantiation serves every nv.
//
// Work items, chunking, fused MSE partials and the per-lane summation order
// (datapoint lane + 64 k, k ascending) are those of k_batched, so MSEs are
// bit-identical to it for equal chunk and queue count.
#pragma once

#include "sr_device.h"
#include "sr_kernels.h"

namespace sr {

constexpr int kVecD = 6;                 // datapoints per lane
constexpr int kVecTile = kWave * kVecD;  // 384 datapoints per wave-instruction
constexpr int kVecLdsSlots = 4;          // stack slots in LDS; deeper ones spill
constexpr int kVecSpillSlots = 8 - kVecLdsSlots;

using ConstIns = const __attribute__((address_space(4))) Ins;

// Instruction i of a wave-uniform program through the constant address
// space, so that it is fetched with a scalar load.
SR_DEV Ins fetch_ins(ConstIns *p, int i)
{
    return Ins{p[i].code, p[i].arg};
}

// A raw buffer resource over `bytes` bytes at p (scalar registers): a load is
// then one buffer instruction from a 32-bit per-lane byte offset, and a load
// at or past `bytes` returns 0 without touching memory (the range check of
// raw buffers, stride 0).  Word 3 is the gfx9-family raw-buffer format word.
SR_DEV __amdgpu_buffer_rsrc_t col_rsrc(const float *p, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, (int)bytes, 0x00020000);
}

template <typename F>
SR_DEV void vec_each(F &&f)
{
#pragma unroll
    for (int k = 0; k < kVecD; ++k) f(k);
}

// The multi-instruction library ops on the datapoints in two halves, with a
// scheduling barrier between them: their temporaries are live for half the
// datapoints at a time (fewer registers at the kernel's peak).
template <class F>
SR_DEV void vec_halves(F &&f)
{
#pragma unroll
    for (int k = 0; k < kVecD / 2; ++k) f(k);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int k = kVecD / 2; k < kVecD; ++k) f(k);
}

SR_DEV void vec_unary(int op, float (&t)[kVecD])
{
    using namespace sr_ops;
    switch (op) {
    case OP_SIN: vec_halves([&](int k) { t[k] = f_sin(t[k]); }); break;
    case OP_COS: vec_halves([&](int k) { t[k] = f_cos(t[k]); }); break;
    case OP_TAN: vec_halves([&](int k) { t[k] = f_tan(t[k]); }); break;
    case OP_SINH: vec_halves([&](int k) { t[k] = f_sinh(t[k]); }); break;
    case OP_COSH: vec_halves([&](int k) { t[k] = f_cosh(t[k]); }); break;
    case OP_TANH: vec_halves([&](int k) { t[k] = f_tanh(t[k]); }); break;
    case OP_EXP: vec_halves([&](int k) { t[k] = f_exp(t[k]); }); break;
    case OP_LOG: vec_halves([&](int k) { t[k] = f_log(t[k]); }); break;
    case OP_INV: vec_each([&](int k) { t[k] = f_inv(t[k]); }); break;
    case OP_ASIN: vec_halves([&](int k) { t[k] = f_asin(t[k]); }); break;
    case OP_ACOS: vec_halves([&](int k) { t[k] = f_acos(t[k]); }); break;
    case OP_ATAN: vec_halves([&](int k) { t[k] = f_atan(t[k]); }); break;
    case OP_LOOSE_LOG: vec_each([&](int k) { t[k] = f_loose_log(t[k]); }); break;
    case OP_LOOSE_INV: vec_each([&](int k) { t[k] = f_loose_inv(t[k]); }); break;
    case OP_ABS: vec_each([&](int k) { t[k] = f_abs(t[k]); }); break;
    case OP_NEG: vec_each([&](int k) { t[k] = f_neg(t[k]); }); break;
    case OP_SQRT: vec_each([&](int k) { t[k] = f_sqrt(t[k]); }); break;
    case OP_LOOSE_SQRT: vec_each([&](int k) { t[k] = f_loose_sqrt(t[k]); }); break;
    case OP_EXP_CLAMPED: vec_each([&](int k) { t[k] = f_exp_clamped(t[k]); }); break;
    case OPC_STORE: break;
    default: vec_each([&](int k) { t[k] = 0.0f; }); break;
    }
}

SR_DEV void vec_binary(int op, float (&t)[kVecD], const float (&o)[kVecD])
{
    using namespace sr_ops;
    switch (op) {
    case OP_ADD: vec_each([&](int k) { t[k] = f_add(t[k], o[k]); }); break;
    case OP_SUB: vec_each([&](int k) { t[k] = f_sub(t[k], o[k]); }); break;
    case OP_MUL: vec_each([&](int k) { t[k] = f_mul(t[k], o[k]); }); break;
    case OP_DIV: vec_each([&](int k) { t[k] = f_div(t[k], o[k]); }); break;
    case OP_POW: vec_each([&](int k) { t[k] = f_pow(t[k], o[k]); }); break;
    case OP_MIN: vec_each([&](int k) { t[k] = f_min(t[k], o[k]); }); break;
    case OP_MAX: vec_each([&](int k) { t[k] = f_max(t[k], o[k]); }); break;
    case OP_LOOSE_DIV: vec_each([&](int k) { t[k] = f_loose_div(t[k], o[k]); }); break;
    case OP_LOOSE_POW: vec_each([&](int k) { t[k] = f_loose_pow(t[k], o[k]); }); break;
    default: vec_each([&](int k) { t[k] = 0.0f; }); break;
    }
}

// Same as vec_binary with the operands exchanged (top = f(o, top)).
SR_DEV void vec_binary_swapped(int op, float (&t)[kVecD], const float (&o)[kVecD])
{
    float a[kVecD];
    vec_each([&](int k) {
        a[k] = o[k];
    });
    vec_binary(op, a, t);
    vec_each([&](int k) { t[k] = a[k]; });
}

template <int MODE>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6))) void k_vec(BatchArgs a)
{
    __shared__ float s_stk[kBlock / kWave][kVecLdsSlots][kVecD][kWave];
    const int w = threadIdx.x / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    float(*stk)[kVecD][kWave] = s_stk[w];
    float *spill = a.spill + ((int64_t)blockIdx.x * (kBlock / kWave) + w) * kVecSpillSlots * kVecTile;
    const int q = a.n_queues > 1 ? (int)(blockIdx.x % (unsigned)a.n_queues) : 0;
    const int n_exprs = a.n_exprs_dev ? min(*a.n_exprs_dev, a.n_exprs) : a.n_exprs;
    const int items = n_exprs * a.chunks_per_queue;
    const int64_t N = a.N;

    for (;;) {
        int item = 0;
        if (lane == 0) item = (int)atomicAdd(&a.counters[q], 1u);
        item = __builtin_amdgcn_readfirstlane(item);
        if (item >= items) break;
        const int chunk = item / n_exprs;
        const int rank = item - chunk * n_exprs;
        // the item's expression, program and length are wave-uniform: held
        // in scalar registers, so each instruction fetch is a scalar load
        // from a scalar base
        const int e = __builtin_amdgcn_readfirstlane(a.order[rank]);
        const int64_t poff =
            a.prog_off ? (int64_t)__builtin_amdgcn_readfirstlane(a.prog_off[e]) : (int64_t)e * a.prog_stride;
        ConstIns *prog = (ConstIns *)(a.progs + poff);
        const int n = __builtin_amdgcn_readfirstlane(a.prog_len[e]);
        const int64_t q0 = (int64_t)q * a.slice;
        const int64_t q1 = q0 + a.slice < N ? q0 + a.slice : N;
        const int64_t c0 = q0 + (int64_t)chunk * a.chunk;
        const int64_t c1 = c0 + a.chunk < q1 ? c0 + a.chunk : q1;
        const bool need = !has_ref(MODE) || !a.need_mse || __builtin_amdgcn_readfirstlane(a.need_mse[e]) != 0;
        // per-lane running sum (fp64, datapoints in ascending order) and
        // count of valid datapoints (at most chunk / 64 per lane)
        double sum = 0.0;
        int cnt = 0;

        for (int64_t t0 = c0; t0 < c1; t0 += kVecTile) {
            const int last = (int)((c1 - t0 < kVecTile ? c1 - t0 : kVecTile) - 1);
            const bool full = last == kVecTile - 1;
            // six coalesced loads per lane from a uniform base (datapoint
            // t0 + 64 k + lane): one per-lane byte offset and a scalar one per
            // k; the resource's record count ends at the tile's last
            // datapoint, so on a ragged tile the loads past it read 0 (those
            // lanes are masked out of the MSE and the column stores)
            const int lane4 = lane * 4;
            const uint32_t tile_bytes = (uint32_t)(last + 1) * 4u;
            auto load = [&](const float *base, float(&dst)[kVecD]) {
                const __amdgpu_buffer_rsrc_t r = col_rsrc(base, tile_bytes);
                vec_each([&](int k) {
                    dst[k] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, lane4, k * kWave * 4, 0));
                });
            };
            // operand of an instruction: its variable / column, a constant,
            // or nothing (a popped operand is read when the instruction runs);
            // column arg starts arg * N elements past the tile's base (an
            // unsigned 32 x 32-bit product: N < 2^32 for this kernel)
            const float *xt = a.X + t0;
            const float *ct = a.cache + t0;
            static_assert(SRC_REF == (SRC_VAR | 1), "variable and column sources differ in bit 0");
            auto in_memory = [&](int code) {
                return has_ref(MODE) ? (bc_src(code) | 1) == SRC_REF : bc_src(code) == SRC_VAR;
            };
            float top[kVecD], o[kVecD];
            // the target, needed after the program (only when this item's MSE
            // is): loaded into the operand buffer of instruction n, the first
            // one no instruction uses, so it arrives during the program's last
            // instructions and takes no registers of its own
            const float *yt = a.X + (int64_t)a.nv * N + t0;
            vec_each([&](int k) { top[k] = 0.0f; });
            int sp = 0;
            // operands two instructions ahead: instruction i reads buffer
            // i % 2, which then receives the operand of instruction i + 2
            float ba[kVecD], bb[kVecD];
            int code_a = 0, arg_a = 0, code_b = 0, arg_b = 0;
            // one test for "operand in memory" (variable or column), then a
            // select of the base: fewer scalar branches per instruction
            auto fetch_to = [&](int i, int &c_, int &a_, float(&B)[kVecD]) {
                if (i < n) {
                    const Ins x = fetch_ins(prog, i);
                    c_ = __builtin_amdgcn_readfirstlane(x.code);
                    a_ = __builtin_amdgcn_readfirstlane(x.arg);
                    if (in_memory(c_)) {
                        const float *base = has_ref(MODE) && bc_src(c_) == SRC_REF ? ct : xt;
                        load(base + (uint64_t)(uint32_t)a_ * (uint32_t)N, B);
                    }
                } else if (i == n && need) {
                    load(yt, B);
                }
            };
            fetch_to(0, code_a, arg_a, ba);
            fetch_to(1, code_b, arg_b, bb);
            auto step = [&](int i, int &code_r, int &arg_r, float(&B)[kVecD]) {
                const int code = code_r, arg = arg_r;
                const int src = bc_src(code);
                if (in_memory(code)) vec_each([&](int k) { o[k] = B[k]; });
                fetch_to(i + 2, code_r, arg_r, B);
                if (code & BC_PUSH) {
                    if (sp < kVecLdsSlots) vec_each([&](int k) { stk[sp][k][lane] = top[k]; });
                    else vec_each([&](int k) { spill[(sp - kVecLdsSlots) * kVecTile + k * kWave + lane] = top[k]; });
                    ++sp;
                }
                if (src == SRC_POP) {
                    --sp;
                    if (sp < kVecLdsSlots) vec_each([&](int k) { o[k] = stk[sp][k][lane]; });
                    else {
                        vec_each([&](int k) { o[k] = spill[(sp - kVecLdsSlots) * kVecTile + k * kWave + lane]; });
                        // settle these loads here: left pending, they would make every
                        // later write of the operand registers, on any path, wait for all
                        // outstanding loads -- the prefetched operands included
                        __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0)
                    }
                } else if (src == SRC_CONST) {
                    const float cv = __int_as_float(arg);
                    vec_each([&](int k) { o[k] = cv; });
                }
                const int cls = bc_cls(code);
                if (cls != CL_BIN) {
                    if (src != SRC_NONE) vec_each([&](int k) { top[k] = o[k]; });
                    if (cls == CL_UN) vec_unary(bc_op(code), top);
                } else if (code & BC_SWAP) {
                    vec_binary_swapped(bc_op(code), top, o);
                } else {
                    vec_binary(bc_op(code), top, o);
                }
                if (has_ref(MODE) && code == STORE_CODE) {
                    float *cs = a.out + (int64_t)arg * N + t0;
                    if (full) vec_each([&](int k) { cs[k * kWave + lane] = top[k]; });
                    else vec_each([&](int k) {
                            const int j = k * kWave + lane;
                            if (j <= last) cs[j] = top[k];
                        });
                }
            };
            for (int i = 0; i < n; i += 2) {
                step(i, code_a, arg_a, ba);
                if (i + 1 < n) step(i + 1, code_b, arg_b, bb);
            }
            const bool y_in_b = n & 1;
            if (need) vec_each([&](int k) {
                const bool ok = k * kWave + lane <= last && isfinite(top[k]);
                const float diff = top[k] - (y_in_b ? bb[k] : ba[k]);
                // an invalid datapoint adds +0.0: the sum is never -0.0, so
                // this leaves it unchanged
                sum += (double)(ok ? diff * diff : 0.0f);
                cnt += ok ? 1 : 0;
            });
        }
        long long wcnt = 0;
        if (need) {
            sum = wave_sum(sum);
            wcnt = wave_sum((long long)cnt);
        }
        if (lane == 0) {
            const int slot = q * a.chunks_per_queue + chunk;
            a.psum[(int64_t)e * a.P + slot] = sum;
            a.pcnt[(int64_t)e * a.P + slot] = wcnt;
        }
    }
}

}  // namespace sr
