// sr_ops.h -- scalar fp32 semantics of every opcode, restating
// cl_cpu_eval.cpp:16-56 (the reference CPU evaluator): DIV by 0 -> NaN,
// LOOSE_* guards with DELTA = 1e-9 (a zero denominator takes +DELTA) and
// MAX_VAL = 1e9, LOOSE_POW(0, 0) = 0, MIN/MAX = fminf/fmaxf, unknown op -> 0.
// Opcode 29 is the clamped exp of evolution_tests.cpp:40.
//
// The interpreter kernels call these functions, and the hipRTC backend embeds
// this file verbatim in every generated module, so both evaluate each node
// with the same code.  Self-contained on purpose: no includes.
#ifndef SR_OPS_H
#define SR_OPS_H

#ifndef SR_OPF
#define SR_OPF __device__ __forceinline__
#endif
// Library-backed (multi-instruction) functions; generated modules may make
// them out-of-line calls to keep code size and compile time down.
#ifndef SR_OPF_T
#define SR_OPF_T SR_OPF
#endif

namespace sr_ops {

constexpr float kDelta = 1e-9f;
constexpr float kMaxVal = 1e9f;

SR_OPF float guard(float b) { return fabsf(b) <= kDelta ? (b < 0.0f ? -kDelta : kDelta) : b; }

SR_OPF float f_add(float a, float b) { return a + b; }
SR_OPF float f_sub(float a, float b) { return a - b; }
SR_OPF float f_mul(float a, float b) { return a * b; }
SR_OPF float f_div(float a, float b) { return b == 0.0f ? __builtin_nanf("") : a / b; }
SR_OPF_T float f_pow(float a, float b) { return powf(a, b); }
SR_OPF float f_min(float a, float b) { return fminf(a, b); }
SR_OPF float f_max(float a, float b) { return fmaxf(a, b); }
SR_OPF float f_loose_div(float a, float b) { return a / guard(b); }
SR_OPF_T float f_loose_pow(float a, float b) { return (a == 0.0f && b == 0.0f) ? 0.0f : powf(fabsf(a), b); }

SR_OPF_T float f_sin(float a) { return sinf(a); }
SR_OPF_T float f_cos(float a) { return cosf(a); }
SR_OPF_T float f_tan(float a) { return tanf(a); }
SR_OPF_T float f_sinh(float a) { return sinhf(a); }
SR_OPF_T float f_cosh(float a) { return coshf(a); }
SR_OPF_T float f_tanh(float a) { return tanhf(a); }
SR_OPF_T float f_exp(float a) { return expf(a); }
SR_OPF_T float f_log(float a) { return logf(a); }
SR_OPF float f_inv(float a) { return 1.0f / a; }
SR_OPF_T float f_asin(float a) { return asinf(a); }
SR_OPF_T float f_acos(float a) { return acosf(a); }
SR_OPF_T float f_atan(float a) { return atanf(a); }
SR_OPF_T float f_loose_log(float a) { return a == 0.0f ? -kMaxVal : logf(fabsf(a)); }
SR_OPF float f_loose_inv(float a) { return 1.0f / guard(a); }
SR_OPF float f_abs(float a) { return fabsf(a); }
SR_OPF float f_neg(float a) { return -a; }
SR_OPF float f_sqrt(float a) { return sqrtf(a); }
SR_OPF float f_loose_sqrt(float a) { return sqrtf(fabsf(a)); }
SR_OPF_T float f_exp_clamped(float a) { return a > 10.0f ? kMaxVal : expf(a); }
SR_OPF float f_zero(float) { return 0.0f; }
SR_OPF float f_zero2(float, float) { return 0.0f; }

}  // namespace sr_ops

#endif
