// sr_device.h -- device-side op semantics and the per-lane interpreter state
// shared by every evaluation kernel (gfx950, wave64).
//
// Op semantics restate cl_cpu_eval.cpp:16-56 (the reference CPU evaluator):
// fp32 arithmetic, DIV by 0 -> NaN, LOOSE_* guards with DELTA = 1e-9 and
// MAX_VAL = 1e9, unknown unary/binary op -> 0.  Opcode 29 is the clamped exp of
// evolution_tests.cpp:40 (a > 10 -> 1e9, else expf(a)).  All code is built
// with -ffp-contract=off so no product/sum is fused across tokens.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sr_format.h"
#include "sr_ops.h"

#define SR_DEV __device__ __forceinline__

namespace sr {

constexpr float kDelta = 1e-9f;
constexpr float kMaxVal = 1e9f;

SR_DEV float qnan() { return __builtin_nanf(""); }

// ---------------------------------------------------------------------------
// Per-datapoint lane state.  Each datapoint a lane evaluates owns its own
// small arrays so the compiler keeps each one in VGPRs and indexes it with
// s_set_gpr_idx (the index -- stack pointer or variable id -- is wave-uniform).
// ---------------------------------------------------------------------------
// Variable registers per datapoint.  The backend reads an array of <= 8
// floats at a uniform dynamic index through a compare/select chain over every
// element, larger ones through GPR-index mode (one move).  With <= 4
// variables the 4-element chain (3 selects; 2 for a 3-element array) costs
// less than the 5 extra registers per datapoint a 9-element array would take
// from occupancy.
template <int XC>
constexpr int kXRegs = XC > 8 ? XC : (XC <= 4 ? XC : 9);

template <int S, int XC>
struct Lane {
    float (&stk)[S];
    float (&x)[kXRegs<XC>];
    float &top;
    float &o;  // second operand / loaded leaf
    float (&xn)[XC];  // next tile's variables (loaded ahead)
    float &y, &yn;    // target of this / the next tile
};

// Apply a unary op (10..27, 29) to every lane's top.  op is wave-uniform.
template <typename... L>
SR_DEV void apply_unary(int op, L &...l)
{
    using namespace sr_ops;
    switch (op) {
    case OP_SIN: ((l.top = f_sin(l.top)), ...); break;
    case OP_COS: ((l.top = f_cos(l.top)), ...); break;
    case OP_TAN: ((l.top = f_tan(l.top)), ...); break;
    case OP_SINH: ((l.top = f_sinh(l.top)), ...); break;
    case OP_COSH: ((l.top = f_cosh(l.top)), ...); break;
    case OP_TANH: ((l.top = f_tanh(l.top)), ...); break;
    case OP_EXP: ((l.top = f_exp(l.top)), ...); break;
    case OP_LOG: ((l.top = f_log(l.top)), ...); break;
    case OP_INV: ((l.top = f_inv(l.top)), ...); break;
    case OP_ASIN: ((l.top = f_asin(l.top)), ...); break;
    case OP_ACOS: ((l.top = f_acos(l.top)), ...); break;
    case OP_ATAN: ((l.top = f_atan(l.top)), ...); break;
    case OP_LOOSE_LOG: ((l.top = f_loose_log(l.top)), ...); break;
    case OP_LOOSE_INV: ((l.top = f_loose_inv(l.top)), ...); break;
    case OP_ABS: ((l.top = f_abs(l.top)), ...); break;
    case OP_NEG: ((l.top = f_neg(l.top)), ...); break;
    case OP_SQRT: ((l.top = f_sqrt(l.top)), ...); break;
    case OP_LOOSE_SQRT: ((l.top = f_loose_sqrt(l.top)), ...); break;
    case OP_EXP_CLAMPED: ((l.top = f_exp_clamped(l.top)), ...); break;
    case OPC_STORE: break;  // the cache store pseudo-op leaves top unchanged
    default: ((l.top = 0.0f), ...); break;
    }
}

// top = f(top, o) for binary ops 1..9.
template <typename... L>
SR_DEV void apply_binary(int op, L &...l)
{
    using namespace sr_ops;
    switch (op) {
    case OP_ADD: ((l.top = f_add(l.top, l.o)), ...); break;
    case OP_SUB: ((l.top = f_sub(l.top, l.o)), ...); break;
    case OP_MUL: ((l.top = f_mul(l.top, l.o)), ...); break;
    case OP_DIV: ((l.top = f_div(l.top, l.o)), ...); break;
    case OP_POW: ((l.top = f_pow(l.top, l.o)), ...); break;
    case OP_MIN: ((l.top = f_min(l.top, l.o)), ...); break;
    case OP_MAX: ((l.top = f_max(l.top, l.o)), ...); break;
    case OP_LOOSE_DIV: ((l.top = f_loose_div(l.top, l.o)), ...); break;
    case OP_LOOSE_POW: ((l.top = f_loose_pow(l.top, l.o)), ...); break;
    default: ((l.top = 0.0f), ...); break;
    }
}

// Recursively declare DPT separate Lane objects (one per datapoint a lane
// owns) and hand them to f as a parameter pack.
template <int DPT, int S, int XC, typename F, typename... L>
SR_DEV void with_lanes(F &&f, L &...ls)
{
    if constexpr (sizeof...(L) == DPT) {
        f(ls...);
    } else {
        // Separate locals per datapoint: each array is its own alloca, which
        // the compiler promotes to VGPRs (a struct of arrays would not be).
        float stk[S];
        float x[kXRegs<XC>], xn[XC];
        float top = 0.0f, o = 0.0f, y = 0.0f, yn = 0.0f;
        Lane<S, XC> l{stk, x, top, o, xn, y, yn};
        with_lanes<DPT, S, XC>(f, ls..., l);
    }
}

// Fold helper with the lane's ordinal: g(d, lane).
template <typename G, typename... L>
SR_DEV void for_each_lane(G &&g, L &...ls)
{
    int d = 0;
    ((g(d++, ls)), ...);
}

// ---------------------------------------------------------------------------
// Wave reductions (wave64, deterministic shuffle-xor tree).
// ---------------------------------------------------------------------------
SR_DEV double wave_sum(double v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

SR_DEV long long wave_sum(long long v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

SR_DEV float wave_sum(float v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

SR_DEV int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

}  // namespace sr
