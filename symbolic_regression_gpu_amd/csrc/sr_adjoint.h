// sr_adjoint.h -- constant fitting: forward evaluation with free constants
// read from a device array (several replicas at once, e.g. central finite
// differences), and a reverse-mode adjoint pass giving dMSE/dconstants for
// every expression in one launch.
//
// Adjoint of one datapoint (DPT = 1 per lane):
//   forward : run the bytecode, tape[i] = top after instruction i
//   seed    : a = pred - y for a finite prediction (its MSE term), else 0
//   reverse : instructions n-1 .. 0; the operands of instruction i are
//             tape[i-1] (previous top) and its source (constant, variable,
//             or tape[arg] for a popped value, the pop's producer index being
//             stored in arg); local partials give the adjoints of top and
//             source; a popped source's adjoint is pushed on an adjoint stack
//             and returned to the instruction that pushed the value.
//   dMSE/dc_k = 2/n_valid * sum over valid points of a * d pred / d c_k.
// The tape lives in a global scratch region per wave (one coalesced 256-byte
// row per instruction, L2 / MALL resident); a dynamically indexed tape does
// not stay in VGPRs under the compiler's alloca promotion rules.
#pragma once

#include "sr_kernels.h"

namespace sr {

struct ParamArgs {
    const Ins *progs;
    const int32_t *aux;       // per instruction, packed like progs (sr_adjoint_prog.h)
    const int32_t *prog_off;
    const int32_t *prog_len;
    const int32_t *order;
    const int32_t *nslots;  // free constants per expression
    const float *X;
    const float *cvals;     // [E][MAX_SLOTS] current constant values
    double *psum;           // [R*E][P]
    long long *pcnt;        // [R*E][P]
    double *pgrad;          // adjoint: [E][P][MAX_SLOTS]
    float *tape;            // global tape scratch (long programs)
    uint32_t *counters;
    int64_t N, slice, chunk;
    int n_exprs, nv, n_queues, chunks_per_queue, P;
    int replicas;           // forward: 1 = plain, 2*MAX_SLOTS = central differences
    float fd_h;
    int tape_stride;        // tape rows per wave
    int prog_lds;           // instructions per wave in the program LDS (longest program + 1)
};

// Dynamic LDS of the parametric kernels: each wave's program (Ins) and, for
// the adjoint, its aux words.
inline size_t param_lds_bytes(int prog_lds, bool adjoint)
{
    return (size_t)(kBlock / kWave) * prog_lds * (sizeof(Ins) + (adjoint ? sizeof(int32_t) : 0));
}

// d f(u) / du for unary op, given input u and output y.
SR_DEV float d_unary(int op, float u, float y)
{
    switch (op) {
    case OP_SIN: return cosf(u);
    case OP_COS: return -sinf(u);
    case OP_TAN: return 1.0f + y * y;
    case OP_SINH: return coshf(u);
    case OP_COSH: return sinhf(u);
    case OP_TANH: return 1.0f - y * y;
    case OP_EXP: return y;
    case OP_LOG: return 1.0f / u;
    case OP_INV: return -(y * y);
    case OP_ASIN: return 1.0f / sqrtf(1.0f - u * u);
    case OP_ACOS: return -1.0f / sqrtf(1.0f - u * u);
    case OP_ATAN: return 1.0f / (1.0f + u * u);
    case OP_LOOSE_LOG: return u == 0.0f ? 0.0f : 1.0f / u;
    case OP_LOOSE_INV: return fabsf(u) <= kDelta ? 0.0f : -(y * y);
    case OP_ABS: return u > 0.0f ? 1.0f : (u < 0.0f ? -1.0f : 0.0f);
    case OP_NEG: return -1.0f;
    case OP_SQRT: return 0.5f / y;
    case OP_LOOSE_SQRT: return u == 0.0f ? 0.0f : (u > 0.0f ? 0.5f : -0.5f) / y;
    case OP_EXP_CLAMPED: return u > 10.0f ? 0.0f : y;
    default: return 0.0f;
    }
}

// Partials of a binary op at (a, b) with output y.
SR_DEV void d_binary(int op, float a, float b, float y, float &da, float &db)
{
    switch (op) {
    case OP_ADD: da = 1.0f; db = 1.0f; break;
    case OP_SUB: da = 1.0f; db = -1.0f; break;
    case OP_MUL: da = b; db = a; break;
    case OP_DIV: da = 1.0f / b; db = -y / b; break;
    case OP_POW:
        da = a == 0.0f ? (b == 1.0f ? 1.0f : 0.0f) : b * y / a;
        db = a > 0.0f ? y * logf(a) : 0.0f;
        break;
    case OP_MIN: {
        const bool sel_a = (b != b) || (a == a && a <= b);
        da = sel_a ? 1.0f : 0.0f;
        db = sel_a ? 0.0f : 1.0f;
        break;
    }
    case OP_MAX: {
        const bool sel_a = (b != b) || (a == a && a >= b);
        da = sel_a ? 1.0f : 0.0f;
        db = sel_a ? 0.0f : 1.0f;
        break;
    }
    case OP_LOOSE_DIV: {
        const bool small = fabsf(b) <= kDelta;
        const float d = small ? (b < 0.0f ? -kDelta : kDelta) : b;
        da = 1.0f / d;
        db = small ? 0.0f : -y / d;
        break;
    }
    case OP_LOOSE_POW:
        if (a == 0.0f) {
            da = 0.0f;
            db = 0.0f;
        } else {
            da = b * y / a;
            db = y * logf(fabsf(a));
        }
        break;
    default: da = 0.0f; db = 0.0f; break;
    }
}

}  // namespace sr
