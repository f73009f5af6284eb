/*
 * sr_oracle.c -- CPU restatement of the reference hot path (test infrastructure).
 *
 * It is the oracle of the HIP product path and its CPU timing baseline.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 * See sr_oracle.h for what pins each function.
 *
 * Build: gcc -O2 -ffp-contract=off -fPIC -shared (oracle/Makefile).  FP
 * contraction is off so every fp32 operation rounds exactly as written.
 */
#include "sr_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Op semantics                                                              */
/* ------------------------------------------------------------------------ */

/* cl_cpu_eval.cpp:16-56 (fp32 op table).  Opcode 29 (exp clamped) follows the
 * clamped exp of evolution_tests.cpp:40: a > 10 -> 1e9, else expf(a). */
float sro_op32(int op, float a, float b)
{
    const float DELTA = 1e-9f;
    const float MAX_VAL = 1e9f;
    switch (op) {
    case 1: return a + b;
    case 2: return a - b;
    case 3: return a * b;
    case 4: return (b == 0.0f) ? NAN : a / b;
    case 5: return powf(a, b);
    case 6: return fminf(a, b);
    case 7: return fmaxf(a, b);
    case 8: {
        float d = fabsf(b) <= DELTA ? (b < 0.0f ? -DELTA : DELTA) : b;
        return a / d;
    }
    case 9: return (a == 0.0f && b == 0.0f) ? 0.0f : powf(fabsf(a), b);
    case 10: return sinf(a);
    case 11: return cosf(a);
    case 12: return tanf(a);
    case 13: return sinhf(a);
    case 14: return coshf(a);
    case 15: return tanhf(a);
    case 16: return expf(a);
    case 17: return logf(a);
    case 18: return 1.0f / a;
    case 19: return asinf(a);
    case 20: return acosf(a);
    case 21: return atanf(a);
    case 22: return (a == 0.0f) ? -MAX_VAL : logf(fabsf(a));
    case 23: {
        float d = fabsf(a) <= DELTA ? (a < 0.0f ? -DELTA : DELTA) : a;
        return 1.0f / d;
    }
    case 24: return fabsf(a);
    case 25: return -a;
    case 26: return sqrtf(a);
    case 27: return sqrtf(fabsf(a));
    case SRO_OP_EXP_CLAMPED: return (a > 10.0f) ? MAX_VAL : expf(a);
    default: return 0.0f;
    }
}

/* cpu_simple_single.cpp:6-79 (fp64 twin; MIN/MAX via <= / >=). */
double sro_op64(int op, double a, double b)
{
    const double DELTA = 1e-9;
    const double MAX_VAL = 1e9;
    switch (op) {
    case 1: return a + b;
    case 2: return a - b;
    case 3: return a * b;
    case 4: return (b == 0.0) ? NAN : a / b;
    case 5: return pow(a, b);
    case 6: return (a <= b) ? a : b;
    case 7: return (a >= b) ? a : b;
    case 8: {
        double d = fabs(b) <= DELTA ? (b < 0 ? -DELTA : DELTA) : b;
        return a / d;
    }
    case 9: return (a == 0.0 && b == 0.0) ? 0.0 : pow(fabs(a), b);
    case 10: return sin(a);
    case 11: return cos(a);
    case 12: return tan(a);
    case 13: return sinh(a);
    case 14: return cosh(a);
    case 15: return tanh(a);
    case 16: return exp(a);
    case 17: return log(a);
    case 18: return 1.0 / a;
    case 19: return asin(a);
    case 20: return acos(a);
    case 21: return atan(a);
    case 22: return (a == 0.0) ? -MAX_VAL : log(fabs(a));
    case 23: {
        double d = fabs(a) <= DELTA ? (a < 0 ? -DELTA : DELTA) : a;
        return 1.0 / d;
    }
    case 24: return fabs(a);
    case 25: return -a;
    case 26: return sqrt(a);
    case 27: return sqrt(fabs(a));
    case SRO_OP_EXP_CLAMPED: return (a > 10.0) ? MAX_VAL : exp(a);
    default: return 0.0;
    }
}

/* cl_cpu_eval.cpp:82 -- tokens 1..9 pop two operands, every other positive
 * token pops one. */
int sro_is_binary(int tok) { return tok >= 1 && tok <= 9; }

/* cl_cpu_eval.cpp:59-90: right-to-left prefix scan. */
float sro_eval_point32(const int *tok, const double *val, int len, const float *x, int nv)
{
    float stk[SRO_MAX_STACK];
    int sp = 0;
    for (int i = len - 1; i >= 0; i--) {
        int t = tok[i];
        if (t == SRO_TOK_CONST) {
            stk[sp++] = (float)val[i];
        } else if (t == SRO_TOK_VAR) {
            int v = (int)val[i];
            stk[sp++] = (v < nv) ? x[v] : 0.0f;
        } else {
            float a = stk[--sp];
            float b = 0.0f;
            if (sro_is_binary(t)) b = stk[--sp];
            stk[sp++] = sro_op32(t, a, b);
        }
    }
    return stk[0];
}

/* cpu_simple_single.cpp:96-127 (fp64 interpreter). */
double sro_eval_point64(const int *tok, const double *val, int len, const double *x, int nv)
{
    double stk[SRO_MAX_STACK];
    int sp = 0;
    for (int i = len - 1; i >= 0; i--) {
        int t = tok[i];
        if (t == SRO_TOK_CONST) {
            stk[sp++] = val[i];
        } else if (t == SRO_TOK_VAR) {
            int v = (int)val[i];
            stk[sp++] = (v < nv) ? x[v] : 0.0;
        } else {
            double a = stk[--sp];
            double b = 0.0;
            if (sro_is_binary(t)) b = stk[--sp];
            stk[sp++] = sro_op64(t, a, b);
        }
    }
    return stk[0];
}

/* ---- FP64 value + running error bound ------------------------------------ */

#define U32 (1.0 / 16777216.0) /* fp32 unit roundoff 2^-24 */
#define FLT_MAXD 3.4028234663852886e38
#define TINY 1.5e-45

/* rounding added by one fp32 operation producing y */
static double round_err(double y, double ulps)
{
    return ulps * 2.0 * U32 * fabs(y) + TINY;
}

/* Clamp to fp32 range semantics; returns 1 if ambiguous (within bound). */
static int to_fp32_range(double *y, double *b)
{
    if (isnan(*y) || isinf(*y)) {
        *b = 0.0;
        return 0;
    }
    double ay = fabs(*y);
    if (ay - *b > FLT_MAXD) {
        *y = *y > 0 ? INFINITY : -INFINITY;
        *b = 0.0;
        return 0;
    }
    if (ay + *b > FLT_MAXD) return 1;
    return 0;
}

/* One node of the bound evaluation: FP64 value *y of op t on (a, b) and the
 * error bound *e of its fp32 evaluation from operand bounds ea, eb; sets
 * *unstable when rounding can flip a branch of the op. */
static void bound_step(int t, double a, double ea, double b, double eb, double *yo, double *eo,
                       int *unst)
{
    const double D = 1e-9;
    int unstable = 0;
    /* fp32 semantics in double: MIN/MAX as fmin/fmax (NaN-ignoring) */
    double y = t == 6 ? fmin(a, b) : (t == 7 ? fmax(a, b) : sro_op64(t, a, b)), e = 0.0;
    switch (t) {
    /* fp32 addition never underflows: a zero sum of exact operands is exact */
    case 1: case 2: e = ea + eb + 0.5 * 2.0 * U32 * fabs(y); break;
    case 3: e = fabs(b) * ea + fabs(a) * eb + ea * eb + round_err(y, 0.5); break;
    case 4:
        if (fabs(b) <= eb && !(b == 0.0 && eb == 0.0)) unstable = 1;
        e = (b != 0.0) ? (ea + fabs(y) * eb) / fabs(b) + round_err(y, 0.5) : 0.0;
        break;
    case 5:
    case 9: {
        double aa = t == 9 ? fabs(a) : a;
        /* sign / zero of the base uncertain, or a negative base with an
         * inexact exponent (NaN vs finite can flip) */
        if (fabs(a) <= ea && !(a == 0.0 && ea == 0.0)) unstable = 1;
        if (t == 5 && a < 0.0 && eb > 0.0) unstable = 1;
        /* an exponent that may round to zero turns any base into 1, and
         * so does a base that may round to 1 under a NaN/inf exponent */
        if (fabs(b) <= eb && !(b == 0.0 && eb == 0.0)) unstable = 1;
        if (!isfinite(b) && ea > 0.0 && fabs(fabs(a) - 1.0) <= ea) unstable = 1;
        double da = aa != 0.0 ? fabs(b * y / aa) : 0.0;
        double db = aa > 0.0 ? fabs(y * log(aa)) : 0.0;
        e = da * ea + db * eb + round_err(y, 2.0);
        break;
    }
    case 6: case 7: e = ea > eb ? ea : eb; break;
    case 8: {
        if (fabs(fabs(b) - D) <= eb + 1e-300) unstable = 1;
        double d = fabs(b) <= D ? D : fabs(b);
        e = (ea + (fabs(b) <= D ? 0.0 : fabs(y) * eb)) / d + round_err(y, 0.5);
        break;
    }
    case 10: case 11: e = ea + round_err(y, 2.0); break;
    case 12: e = (1.0 + y * y) * ea + round_err(y, 2.0); break;
    case 13: case 14: e = cosh(a) * ea + round_err(y, 2.0); break;
    case 15: e = (1.0 - y * y) * ea + round_err(y, 2.0); break;
    case 16: e = fabs(y) * ea + round_err(y, 2.0); break;
    case 17:
        if (fabs(a) <= ea && !(a == 0.0 && ea == 0.0)) unstable = 1;
        e = ea / fabs(a) + round_err(y, 2.0);
        break;
    case 18: e = y * y * ea + round_err(y, 0.5); break;
    case 19: case 20:
        if (ea > 0.0 && fabs(fabs(a) - 1.0) <= ea) unstable = 1;
        e = ea / sqrt(fabs(1.0 - a * a)) + round_err(y, 2.0);
        break;
    case 21: e = ea / (1.0 + a * a) + round_err(y, 2.0); break;
    case 22:
        if (fabs(a) <= ea && !(a == 0.0 && ea == 0.0)) unstable = 1;
        e = a != 0.0 ? ea / fabs(a) + round_err(y, 2.0) : 0.0;
        break;
    case 23:
        if (fabs(fabs(a) - D) <= ea + 1e-300) unstable = 1;
        e = (fabs(a) <= D ? 0.0 : y * y * ea) + round_err(y, 0.5);
        break;
    case 24: case 25: e = ea; break;
    case 26:
        if (fabs(a) <= ea && !(a == 0.0 && ea == 0.0)) unstable = 1;
        e = (a > 0.0 ? ea / (2.0 * y) : 0.0) + round_err(y, 0.5);
        break;
    case 27:
        if (fabs(a) <= ea && !(a == 0.0 && ea == 0.0)) unstable = 1;
        e = (a != 0.0 ? ea / (2.0 * y) : 0.0) + round_err(y, 0.5);
        break;
    case SRO_OP_EXP_CLAMPED:
        if (ea > 0.0 && fabs(a - 10.0) <= ea) unstable = 1;
        e = a > 10.0 ? 0.0 : fabs(y) * ea + round_err(y, 2.0);
        break;
    default: e = 0.0; break;
    }
    /* correctly rounded fp32 ops on exact operands whose exact result is
     * an fp32 number return it exactly */
    if (ea == 0.0 && eb == 0.0 && isfinite(y) && (double)(float)y == y &&
        (t == 1 || t == 2 || t == 3 || t == 4 || t == 6 || t == 7 || t == 18 || t == 24 ||
         t == 25 || t == 26 || t == 27 || ((t == 8 || t == 23) && fabs(t == 8 ? b : a) > D)))
        e = 0.0;
    /* first-order propagation only holds while the operand's error is
     * small against the scale on which the op's derivative changes */
    double scale = INFINITY;
    switch (t) {
    case 4: case 8: case 5: case 9: scale = t >= 8 || t == 4 ? fabs(b) : fabs(a); break;
    case 12: scale = fabs(cos(a)); break;
    case 13: case 14: case 16: case SRO_OP_EXP_CLAMPED: scale = 1.0; break;
    case 17: case 18: case 22: case 23: case 26: case 27: scale = fabs(a); break;
    case 19: case 20: scale = 1.0 - fabs(a); break;
    default: break;
    }
    if (t == 5 || t == 9) {
        if (eb > 1e-2 * (1.0 + fabs(b))) unstable = 1;
    }
    if (isfinite(scale) && (t == 4 || t == 8 ? eb : ea) > 1e-2 * scale) unstable = 1;
    /* an infinite FP64 value from inexact operands (pow(1 - d, huge),
     * exp of a huge argument) need not be infinite in fp32 */
    if (isinf(y) && (ea > 0.0 || eb > 0.0)) unstable = 1;
    if (isnan(e)) e = INFINITY;
    if (to_fp32_range(&y, &e)) unstable = 1;
    *yo = y;
    *eo = e;
    if (unstable) *unst = 1;
}

static void eval_bound_point(const int *tok, const double *val, int len, const double *x, int nv,
                             double *out, double *bout)
{
    double sv[SRO_MAX_STACK] = {0}, sb[SRO_MAX_STACK] = {0};
    int sp = 0, unstable = 0;
    for (int i = len - 1; i >= 0; i--) {
        int t = tok[i];
        if (t == SRO_TOK_CONST) {
            sv[sp] = (double)(float)val[i];
            sb[sp++] = 0.0;
            continue;
        }
        if (t == SRO_TOK_VAR) {
            int v = (int)val[i];
            sv[sp] = v < nv ? x[v] : 0.0;
            sb[sp++] = 0.0;
            continue;
        }
        double a = sv[--sp], ea = sb[sp];
        double b = 0.0, eb = 0.0;
        if (sro_is_binary(t)) {
            b = sv[--sp];
            eb = sb[sp];
        }
        double y, e;
        bound_step(t, a, ea, b, eb, &y, &e, &unstable);
        sv[sp] = y;
        sb[sp++] = e;
    }
    *out = sv[0];
    *bout = unstable ? INFINITY : sb[0];
}

/* ---- FP64 gradient of the MSE + error bound of an fp32 adjoint ------------ */

#define RND(p, ulps) ((ulps) * 2.0 * U32 * fabs(p) + TINY)

/* Partials pa = dy/da, pb = dy/db of op t at (a, b) (y its value) as the fp32
 * adjoint pass defines them, and bounds epa, epb on their fp32 evaluation from
 * the operand bounds (second derivative x operand error + rounding).  Returns
 * 1 when rounding can flip which branch of a partial applies. */
static int partials(int t, double a, double ea, double b, double eb, double y, double *pa, double *epa,
                    double *pb, double *epb)
{
    const double D = 1e-9;
    double p = 0.0, q = 0.0, ep = 0.0, eq = 0.0;
    int unst = 0;
    switch (t) {
    case 1: p = 1.0; q = 1.0; break;
    case 2: p = 1.0; q = -1.0; break;
    case 3: p = b; ep = eb; q = a; eq = ea; break;
    case 4:
    case 8:
        if (t == 8 && fabs(b) <= D) {
            p = b < 0.0 ? -1.0 / D : 1.0 / D;
            q = 0.0;
            break;
        }
        p = 1.0 / b;
        ep = eb / (b * b) + RND(p, 0.5);
        q = -a / (b * b);
        eq = ea / (b * b) + 2.0 * fabs(a) * eb / fabs(b * b * b) + RND(q, 1.5);
        break;
    case 5:
    case 9: {
        double aa = t == 9 ? fabs(a) : a, sg = (t == 9 && a < 0.0) ? -1.0 : 1.0;
        if (a == 0.0) {
            /* the kernel's convention at a zero base */
            p = (t == 5 && b == 1.0) ? 1.0 : 0.0;
            q = 0.0;
            break;
        }
        p = b * y / a;
        double lg = aa > 0.0 ? log(aa) : 0.0;
        double dpa_a = b * (b - 1.0) * y / (a * a), dpa_b = (y / a) * (1.0 + b * lg);
        ep = fabs(dpa_a) * ea + fabs(dpa_b) * eb + RND(p, 3.0);
        if (aa > 0.0) {
            q = y * lg;
            double dq_a = sg * (y * b * lg / aa + y / aa), dq_b = y * lg * lg;
            eq = fabs(dq_a) * ea + fabs(dq_b) * eb + RND(q, 3.0);
        }
        break;
    }
    case 6:
    case 7: {
        int sel_a = isnan(b) || (!isnan(a) && (t == 6 ? a <= b : a >= b));
        if (fabs(a - b) <= ea + eb && (ea + eb) > 0.0) unst = 1;
        p = sel_a ? 1.0 : 0.0;
        q = sel_a ? 0.0 : 1.0;
        break;
    }
    case 10: p = cos(a); ep = fmin(ea, 2.0) + RND(p, 2.0); break;
    case 11: p = -sin(a); ep = fmin(ea, 2.0) + RND(p, 2.0); break;
    case 12: p = 1.0 + y * y; ep = fabs(2.0 * y * p) * ea + RND(p, 3.0); break;
    case 13: p = cosh(a); ep = fabs(sinh(a)) * ea + RND(p, 2.0); break;
    case 14: p = sinh(a); ep = cosh(a) * ea + RND(p, 2.0); break;
    case 15: p = 1.0 - y * y; ep = fabs(2.0 * y * p) * ea + RND(p, 3.0) + 4.0 * U32 * y * y; break;
    case 16: p = y; ep = fabs(y) * ea + RND(p, 2.0); break;
    case 17:
    case 22:
        if (a == 0.0) break;
        p = 1.0 / a;
        ep = ea / (a * a) + RND(p, 0.5);
        break;
    case 18: p = -y * y; ep = 2.0 * fabs(y * y * y) * ea + RND(p, 1.5); break;
    case 23:
        if (fabs(a) <= D) break;
        p = -y * y;
        ep = 2.0 * fabs(y * y * y) * ea + RND(p, 1.5);
        break;
    case 19:
    case 20: {
        double w = 1.0 - a * a;
        p = (t == 19 ? 1.0 : -1.0) / sqrt(w);
        ep = fabs(a) / (w * sqrt(w)) * ea + RND(p, 3.0);
        break;
    }
    case 21: {
        double w = 1.0 + a * a;
        p = 1.0 / w;
        ep = 2.0 * fabs(a) / (w * w) * ea + RND(p, 2.0);
        break;
    }
    case 24:
        if (fabs(a) <= ea && ea > 0.0) unst = 1;
        p = a > 0.0 ? 1.0 : (a < 0.0 ? -1.0 : 0.0);
        break;
    case 25: p = -1.0; break;
    case 26:
        if (a <= 0.0) break;
        p = 0.5 / y;
        ep = 0.25 / (a * y) * ea + RND(p, 1.5);
        break;
    case 27:
        if (a == 0.0) break;
        p = (a > 0.0 ? 0.5 : -0.5) / y;
        ep = 0.25 / (fabs(a) * y) * ea + RND(p, 1.5);
        break;
    case SRO_OP_EXP_CLAMPED:
        if (a > 10.0) break;
        p = y;
        ep = fabs(y) * ea + RND(p, 2.0);
        break;
    default: break;
    }
    *pa = p;
    *epa = ep;
    *pb = q;
    *epb = eq;
    return unst;
}

void sro_mse_grad64_bound(const int *tok, const double *val, const int *slot, int len, const float *X,
                          int nv, int64_t N, const float *cvals, int lane_terms, double *mse_out,
                          double *grad_out, double *bound_out, int64_t *count_out)
{
    enum { K = 8 };
    double *v32 = (double *)malloc(sizeof(double) * (size_t)len);
    for (int i = 0; i < len; i++)
        v32[i] = (slot[i] >= 0 && slot[i] < K) ? (double)cvals[slot[i]] : val[i];
    double sv[SRO_MAX_STACK], sb[SRO_MAX_STACK], sd[SRO_MAX_STACK][K], se[SRO_MAX_STACK][K];
    double gsum[K] = {0}, gerr[K] = {0}, gabs[K] = {0}, msum = 0.0;
    int64_t cnt = 0;
    int unstable_any = 0;
    float xf[64];
    double xd[64];
    for (int64_t dp = 0; dp < N; dp++) {
        for (int v = 0; v < nv && v < 64; v++) {
            xf[v] = X[(int64_t)v * N + dp];
            xd[v] = (double)xf[v];
        }
        float p32 = sro_eval_point32(tok, v32, len, xf, nv);
        if (!isfinite(p32)) continue;
        int sp = 0, unst = 0;
        for (int i = len - 1; i >= 0; i--) {
            int t = tok[i];
            if (t == SRO_TOK_CONST || t == SRO_TOK_VAR) {
                int vv = (int)v32[i];
                sv[sp] = t == SRO_TOK_CONST ? (double)(float)v32[i] : (vv < nv ? xd[vv] : 0.0);
                sb[sp] = 0.0;
                for (int k = 0; k < K; k++) {
                    sd[sp][k] = (t == SRO_TOK_CONST && slot[i] == k) ? 1.0 : 0.0;
                    se[sp][k] = 0.0;
                }
                sp++;
                continue;
            }
            int ia = --sp, ib = -1;
            double a = sv[ia], ea = sb[ia], b = 0.0, eb = 0.0;
            if (sro_is_binary(t)) {
                ib = --sp;
                b = sv[ib];
                eb = sb[ib];
            }
            double y, e, pa, epa, pb, epb;
            bound_step(t, a, ea, b, eb, &y, &e, &unst);
            unst |= partials(t, a, ea, b, eb, y, &pa, &epa, &pb, &epb);
            double nd[K], ne[K];
            for (int k = 0; k < K; k++) {
                double ta = pa * sd[ia][k], tb = ib >= 0 ? pb * sd[ib][k] : 0.0;
                nd[k] = ta + tb;
                ne[k] = fabs(pa) * se[ia][k] + epa * fabs(sd[ia][k]) + (ta != 0.0 ? RND(ta, 0.5) : 0.0);
                if (ib >= 0)
                    ne[k] += fabs(pb) * se[ib][k] + epb * fabs(sd[ib][k]) + (tb != 0.0 ? RND(tb, 0.5) : 0.0) +
                             (ta != 0.0 && tb != 0.0 ? RND(nd[k], 0.5) : 0.0);
                if (!isfinite(ne[k])) unst = 1;
            }
            sv[sp] = y;
            sb[sp] = e;
            memcpy(sd[sp], nd, sizeof nd);
            memcpy(se[sp], ne, sizeof ne);
            sp++;
        }
        if (unst || !isfinite(sv[0]) || !isfinite(sb[0])) unstable_any = 1;
        double tgt = (double)X[(int64_t)nv * N + dp];
        double diff = sv[0] - tgt, ediff = sb[0] + RND(diff, 0.5);
        msum += diff * diff;
        cnt++;
        for (int k = 0; k < K; k++) {
            double c = diff * sd[0][k];
            gsum[k] += c;
            gabs[k] += fabs(c);
            gerr[k] += fabs(diff) * se[0][k] + ediff * fabs(sd[0][k]) + (c != 0.0 ? RND(c, 0.5) : 0.0);
        }
    }
    free(v32);
    *count_out = cnt;
    *mse_out = cnt ? msum / (double)cnt : NAN;
    for (int k = 0; k < K; k++) {
        grad_out[k] = cnt ? 2.0 * gsum[k] / (double)cnt : NAN;
        double sum_err = (double)(lane_terms > 1 ? lane_terms : 1) * 1.01 * U32 * gabs[k];
        bound_out[k] = (!cnt || unstable_any) ? INFINITY : 2.0 * (gerr[k] + sum_err) / (double)cnt;
    }
}

void sro_predict64_bound(const int *tok, const double *val, int len, const float *X, int nv,
                         int64_t N, double *out, double *bound_out)
{
    double x[64];
    for (int64_t dp = 0; dp < N; dp++) {
        for (int v = 0; v < nv && v < 64; v++) x[v] = (double)X[(int64_t)v * N + dp];
        eval_bound_point(tok, val, len, x, nv, &out[dp], &bound_out[dp]);
    }
}

int sro_stack_depth(const int *tok, int len)
{
    int sp = 0, mx = 0;
    for (int i = len - 1; i >= 0; i--) {
        int t = tok[i];
        if (t == SRO_TOK_CONST || t == SRO_TOK_VAR || t == SRO_OP_REF) {
            sp++;
        } else if (t > 0) {
            int need = sro_is_binary(t) ? 2 : 1;
            if (sp < need) return -1;
            sp -= need - 1;
        } else {
            return -1;
        }
        if (sp > mx) mx = sp;
    }
    return sp == 1 ? mx : -1;
}

void sro_predict32(const int *tok, const double *val, int len, const float *X, int nv,
                   int64_t N, float *out)
{
    float x[64];
    for (int64_t dp = 0; dp < N; dp++) {
        for (int v = 0; v < nv && v < 64; v++) x[v] = X[(int64_t)v * N + dp];
        out[dp] = sro_eval_point32(tok, val, len, x, nv);
    }
}

void sro_predict64(const int *tok, const double *val, int len, const float *X, int nv,
                   int64_t N, double *out)
{
    double x[64];
    for (int64_t dp = 0; dp < N; dp++) {
        for (int v = 0; v < nv && v < 64; v++) x[v] = (double)X[(int64_t)v * N + dp];
        out[dp] = sro_eval_point64(tok, val, len, x, nv);
    }
}

/* ---- MSE over a population, threads statically partitioned over exprs ---- */

typedef struct {
    const int *tok;
    const double *val;
    const int *off;
    const int *len;
    int e0, e1;
    const float *X;
    int nv;
    int64_t N;
    float *mse32;
    int64_t *valid;
    double *mse64;
} mse_job;

/* cl_cpu_eval.cpp:117-158 */
static void *mse32_worker(void *arg)
{
    mse_job *j = (mse_job *)arg;
    float x[64];
    const float *y = j->X + (int64_t)j->nv * j->N;
    for (int e = j->e0; e < j->e1; e++) {
        const int *t = j->tok + j->off[e];
        const double *v = j->val + j->off[e];
        int L = j->len[e];
        double total = 0.0;
        int64_t valid = 0;
        for (int64_t dp = 0; dp < j->N; dp++) {
            for (int k = 0; k < j->nv && k < 64; k++) x[k] = j->X[(int64_t)k * j->N + dp];
            float pred = sro_eval_point32(t, v, L, x, j->nv);
            if (!isnan(pred) && !isinf(pred)) {
                float diff = pred - y[dp];
                float sq = diff * diff;
                total += sq;
                valid++;
            }
        }
        j->mse32[e] = valid > 0 ? (float)(total / (double)valid) : NAN;
        if (j->valid) j->valid[e] = valid;
    }
    return NULL;
}

static void *mse64_worker(void *arg)
{
    mse_job *j = (mse_job *)arg;
    double x[64];
    const float *y = j->X + (int64_t)j->nv * j->N;
    for (int e = j->e0; e < j->e1; e++) {
        const int *t = j->tok + j->off[e];
        const double *v = j->val + j->off[e];
        int L = j->len[e];
        double total = 0.0;
        int64_t valid = 0;
        for (int64_t dp = 0; dp < j->N; dp++) {
            for (int k = 0; k < j->nv && k < 64; k++) x[k] = (double)j->X[(int64_t)k * j->N + dp];
            double pred = sro_eval_point64(t, v, L, x, j->nv);
            if (!isnan(pred) && !isinf(pred)) {
                double diff = pred - (double)y[dp];
                total += diff * diff;
                valid++;
            }
        }
        j->mse64[e] = valid > 0 ? total / (double)valid : NAN;
    }
    return NULL;
}

static void run_mse(const int *tok, const double *val, const int *off, const int *len, int E,
                    const float *X, int nv, int64_t N, int threads, float *m32,
                    int64_t *valid, double *m64)
{
    if (threads < 1) threads = 1;
    if (threads > E) threads = E > 0 ? E : 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * threads);
    mse_job *jobs = (mse_job *)malloc(sizeof(mse_job) * threads);
    for (int t = 0; t < threads; t++) {
        mse_job *j = &jobs[t];
        j->tok = tok; j->val = val; j->off = off; j->len = len;
        j->e0 = (int)((int64_t)E * t / threads);
        j->e1 = (int)((int64_t)E * (t + 1) / threads);
        j->X = X; j->nv = nv; j->N = N;
        j->mse32 = m32; j->valid = valid; j->mse64 = m64;
        pthread_create(&th[t], NULL, m32 ? mse32_worker : mse64_worker, j);
    }
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    free(th);
    free(jobs);
}

void sro_mse32(const int *tok, const double *val, const int *off, const int *len, int E,
               const float *X, int nv, int64_t N, int threads, float *mse_out,
               int64_t *valid_out)
{
    run_mse(tok, val, off, len, E, X, nv, N, threads, mse_out, valid_out, NULL);
}

void sro_mse64(const int *tok, const double *val, const int *off, const int *len, int E,
               const float *X, int nv, int64_t N, int threads, double *mse_out)
{
    run_mse(tok, val, off, len, E, X, nv, N, threads, NULL, NULL, mse_out);
}

/* ------------------------------------------------------------------------ */
/* Hashing, detection, cache marking                                         */
/* ------------------------------------------------------------------------ */

#define FNV_OFFSET 14695981039346656037ULL
#define FNV_PRIME 1099511628211ULL

static int detect_arity(int t) { return t < 10 ? 2 : 1; } /* detect.cpp:16-19 */

/* detect.cpp:23-47 (recursive size; terminals are tokens <= 0). */
int sro_subtree_size(const int *tok, int start, int total)
{
    if (start >= total) return 0;
    int t = tok[start];
    if (t <= 0) return 1;
    int ar = detect_arity(t), size = 1, pos = start + 1;
    for (int i = 0; i < ar; i++) {
        if (pos >= total) break;
        int s = sro_subtree_size(tok, pos, total);
        size += s;
        pos += s;
    }
    return size;
}

static uint64_t dbl_bits(double d)
{
    uint64_t u;
    memcpy(&u, &d, sizeof u);
    return u;
}

/* detect.cpp:50-68 == cl_gpu_eval.cu:259-272.  The token is widened as a
 * signed int, so -1 hashes as 0xffffffffffffffff. */
uint64_t sro_hash_flat(const int *tok, const double *val, int len)
{
    uint64_t h = FNV_OFFSET;
    for (int i = 0; i < len; i++) {
        h ^= (uint64_t)(int64_t)tok[i];
        h *= FNV_PRIME;
        h ^= dbl_bits(val[i]);
        h *= FNV_PRIME;
    }
    return h;
}

/* detect.cpp:97-118 and the reverse pass at :170-191. */
void sro_subtree_merkle(const int *tok, const double *val, int len, uint64_t *hash_out,
                        int *size_out)
{
    for (int i = len - 1; i >= 0; i--) {
        int t = tok[i];
        uint64_t h = FNV_OFFSET;
        h ^= (uint64_t)(int64_t)t;
        h *= FNV_PRIME;
        h ^= dbl_bits(val[i]);
        h *= FNV_PRIME;
        int size = 1;
        if (t > 0) {
            int pos = i + 1;
            for (int k = 0; k < detect_arity(t); k++) {
                if (pos >= len) break;
                uint64_t ch = hash_out[pos];
                h = h ^ (ch + 0x9e3779b97f4a7c15ULL + (h << 6) + (h >> 2));
                size += size_out[pos];
                pos += size_out[pos];
            }
        }
        hash_out[i] = h;
        size_out[i] = size;
    }
}

typedef struct {
    uint64_t hash;
    int expr, start, size;
} cand_t;

static int cand_cmp(const void *a, const void *b)
{
    const cand_t *x = (const cand_t *)a, *y = (const cand_t *)b;
    if (x->hash != y->hash) return x->hash < y->hash ? -1 : 1;
    if (x->expr != y->expr) return x->expr < y->expr ? -1 : 1;
    return (x->start > y->start) - (x->start < y->start);
}

typedef struct {
    int first, count, size; /* into verified-candidate array */
    uint64_t hash;
} group_t;

static int group_cmp(const void *a, const void *b)
{
    const group_t *x = (const group_t *)a, *y = (const group_t *)b;
    if (x->size != y->size) return x->size > y->size ? -1 : 1;
    if (x->hash != y->hash) return x->hash < y->hash ? -1 : 1;
    return 0;
}

/* detect_common_subtrees, detect.cpp:120-387. */
void sro_detect(const int *tok, const double *val, const int *off, const int *len, int E,
                int min_size, int min_freq, sro_detect_result *res)
{
    memset(res, 0, sizeof *res);
    int64_t T = 0;
    int maxlen = 0;
    for (int e = 0; e < E; e++) {
        T += len[e];
        if (len[e] > maxlen) maxlen = len[e];
    }
    res->hints = (int *)malloc(sizeof(int) * (T > 0 ? T : 1));
    for (int64_t i = 0; i < T; i++) res->hints[i] = -1;

    /* Step 1: reverse pass, candidates of size >= min_size. */
    cand_t *cands = (cand_t *)malloc(sizeof(cand_t) * (T > 0 ? T : 1));
    int64_t nc = 0;
    uint64_t *hs = (uint64_t *)malloc(sizeof(uint64_t) * (maxlen > 0 ? maxlen : 1));
    int *sz = (int *)malloc(sizeof(int) * (maxlen > 0 ? maxlen : 1));
    for (int e = 0; e < E; e++) {
        sro_subtree_merkle(tok + off[e], val + off[e], len[e], hs, sz);
        for (int i = len[e] - 1; i >= 0; i--) {
            if (sz[i] >= min_size) {
                cands[nc].hash = hs[i];
                cands[nc].expr = e;
                cands[nc].start = i;
                cands[nc].size = sz[i];
                nc++;
            }
        }
    }
    free(hs);
    free(sz);
    qsort(cands, (size_t)nc, sizeof(cand_t), cand_cmp);

    /* Step 2: group by hash, verify against the group's first member. */
    cand_t *ver = (cand_t *)malloc(sizeof(cand_t) * (nc > 0 ? nc : 1));
    int64_t nver = 0;
    group_t *groups = (group_t *)malloc(sizeof(group_t) * (nc > 0 ? nc : 1));
    int ng = 0;
    int64_t i = 0;
    while (i < nc) {
        int64_t j = i + 1;
        while (j < nc && cands[j].hash == cands[i].hash) j++;
        if (j - i >= min_freq) {
            const cand_t *r = &cands[i];
            const int *rt = tok + off[r->expr] + r->start;
            const double *rv = val + off[r->expr] + r->start;
            int64_t first = nver;
            for (int64_t k = i; k < j; k++) {
                const cand_t *c = &cands[k];
                if (c->size != r->size) continue;
                const int *ct = tok + off[c->expr] + c->start;
                const double *cv = val + off[c->expr] + c->start;
                int same = 1;
                for (int q = 0; q < r->size; q++) {
                    if (ct[q] != rt[q] || cv[q] != rv[q]) { same = 0; break; }
                }
                if (same) ver[nver++] = *c;
            }
            if (nver - first >= min_freq) {
                groups[ng].first = (int)first;
                groups[ng].count = (int)(nver - first);
                groups[ng].size = r->size;
                groups[ng].hash = r->hash;
                ng++;
            } else {
                nver = first;
            }
        }
        i = j;
    }

    /* Step 3: size descending. */
    qsort(groups, (size_t)ng, sizeof(group_t), group_cmp);

    /* Step 4: result + greedy non-overlapping marking. */
    res->num_subs = ng;
    res->sub_expr = (int *)malloc(sizeof(int) * (ng > 0 ? ng : 1));
    res->sub_start = (int *)malloc(sizeof(int) * (ng > 0 ? ng : 1));
    res->sub_size = (int *)malloc(sizeof(int) * (ng > 0 ? ng : 1));
    res->sub_hash = (uint64_t *)malloc(sizeof(uint64_t) * (ng > 0 ? ng : 1));
    res->occ_offset = (int *)malloc(sizeof(int) * (ng + 1));
    res->occ_expr = (int *)malloc(sizeof(int) * (nver > 0 ? nver : 1));
    res->occ_start = (int *)malloc(sizeof(int) * (nver > 0 ? nver : 1));
    int o = 0;
    int64_t size_sum = 0;
    for (int s = 0; s < ng; s++) {
        const group_t *g = &groups[s];
        const cand_t *r = &ver[g->first];
        res->sub_expr[s] = r->expr;
        res->sub_start[s] = r->start;
        res->sub_size[s] = r->size;
        res->sub_hash[s] = g->hash;
        size_sum += r->size;
        res->occ_offset[s] = o;
        for (int k = 0; k < g->count; k++) {
            const cand_t *c = &ver[g->first + k];
            res->occ_expr[o] = c->expr;
            res->occ_start[o] = c->start;
            o++;
            int *h = res->hints + off[c->expr] + c->start;
            int overlap = 0;
            for (int q = 0; q < r->size; q++) {
                if (h[q] != -1) { overlap = 1; break; }
            }
            if (!overlap) {
                h[0] = s;
                for (int q = 1; q < r->size; q++) h[q] = -2;
            }
        }
    }
    res->occ_offset[ng] = o;
    int64_t covered = 0;
    for (int64_t q = 0; q < T; q++) covered += res->hints[q] != -1;
    res->covered_tokens = covered;
    res->avg_subtree_size = ng > 0 ? (double)size_sum / ng : 0.0;
    res->coverage = T > 0 ? (double)covered / (double)T : 0.0;
    free(cands);
    free(ver);
    free(groups);
}

void sro_detect_free(sro_detect_result *res)
{
    free(res->sub_expr);
    free(res->sub_start);
    free(res->sub_size);
    free(res->sub_hash);
    free(res->occ_offset);
    free(res->occ_expr);
    free(res->occ_start);
    free(res->hints);
    memset(res, 0, sizeof *res);
}

struct sro_cache_keys {
    int n, capacity;
    uint64_t *key;
    int *off, *len; /* body of slot s: tok/val[off[s] .. off[s] + len[s]) */
    int *tok;
    double *val;
    int ntok, captok;
};

sro_cache_keys *sro_cache_keys_create(int capacity)
{
    sro_cache_keys *k = (sro_cache_keys *)calloc(1, sizeof(sro_cache_keys));
    k->capacity = capacity > 0 ? capacity : 0;
    const int c = k->capacity > 0 ? k->capacity : 1;
    k->key = (uint64_t *)malloc(sizeof(uint64_t) * c);
    k->off = (int *)malloc(sizeof(int) * c);
    k->len = (int *)malloc(sizeof(int) * c);
    return k;
}

void sro_cache_keys_free(sro_cache_keys *k)
{
    if (!k) return;
    free(k->key);
    free(k->off);
    free(k->len);
    free(k->tok);
    free(k->val);
    free(k);
}

int sro_cache_keys_get(const sro_cache_keys *k, uint64_t *keys_out)
{
    if (keys_out) memcpy(keys_out, k->key, sizeof(uint64_t) * (size_t)k->n);
    return k->n;
}

/* Slot whose flat hash AND body equal this subtree's, or -1. */
static int find_key(const sro_cache_keys *k, uint64_t h, const int *t, const double *v, int size)
{
    for (int i = 0; i < k->n; i++)
        if (k->key[i] == h && k->len[i] == size && memcmp(k->tok + k->off[i], t, sizeof(int) * size) == 0 &&
            memcmp(k->val + k->off[i], v, sizeof(double) * size) == 0)
            return i;
    return -1;
}

/* cl_gpu_eval.cu:384-397 */
int sro_cache_update(const sro_detect_result *det, const int *tok, const double *val,
                     const int *off, sro_cache_keys *k, int *new_sub, int *new_slot)
{
    int nn = 0;
    for (int s = 0; s < det->num_subs; s++) {
        const int *t = tok + off[det->sub_expr[s]] + det->sub_start[s];
        const double *v = val + off[det->sub_expr[s]] + det->sub_start[s];
        const int size = det->sub_size[s];
        uint64_t h = sro_hash_flat(t, v, size);
        if (find_key(k, h, t, v, size) >= 0) continue;
        if (k->n >= k->capacity) continue;
        int slot = k->n++;
        k->key[slot] = h;
        if (k->ntok + size > k->captok) {
            k->captok = 2 * (k->ntok + size);
            k->tok = (int *)realloc(k->tok, sizeof(int) * (size_t)k->captok);
            k->val = (double *)realloc(k->val, sizeof(double) * (size_t)k->captok);
        }
        k->off[slot] = k->ntok;
        k->len[slot] = size;
        memcpy(k->tok + k->ntok, t, sizeof(int) * size);
        memcpy(k->val + k->ntok, v, sizeof(double) * size);
        k->ntok += size;
        new_sub[nn] = s;
        new_slot[nn] = slot;
        nn++;
    }
    return nn;
}

/* cl_gpu_eval.cu:279-292 (size with arity 2 for tokens < 10, 1 otherwise). */
void sro_mark_cached(const int *tok, const double *val, const int *off, const int *len,
                     int E, const sro_cache_keys *k, int min_size, int *hints)
{
    for (int e = 0; e < E; e++) {
        const int *t = tok + off[e];
        const double *v = val + off[e];
        int L = len[e];
        int *h = hints + off[e];
        for (int i = 0; i < L; i++) h[i] = -1;
        for (int pos = 0; pos < L; pos++) {
            if (h[pos] != -1) continue;
            int size = sro_subtree_size(t, pos, L);
            if (size < min_size) continue;
            if (pos + size > L) continue;
            int slot = find_key(k, sro_hash_flat(t + pos, v + pos, size), t + pos, v + pos, size);
            if (slot < 0) continue;
            int overlap = 0;
            for (int k = 0; k < size; k++)
                if (h[pos + k] != -1) { overlap = 1; break; }
            if (overlap) continue;
            h[pos] = slot;
            for (int k = 1; k < size && pos + k < L; k++) h[pos + k] = -2;
        }
    }
}

/* cl_gpu_eval.cu:433-459 */
int sro_skeleton(const int *tok, const double *val, const int *off, const int *len, int E,
                 const int *hints, int *sk_tok, double *sk_val, int *sk_off, int *sk_len)
{
    int n = 0;
    for (int e = 0; e < E; e++) {
        sk_off[e] = n;
        const int *t = tok + off[e];
        const double *v = val + off[e];
        const int *h = hints + off[e];
        for (int k = 0; k < len[e]; k++) {
            if (h[k] >= 0) {
                int s = sro_subtree_size(t, k, len[e]);
                sk_tok[n] = SRO_OP_REF;
                sk_val[n] = (double)(float)h[k];
                n++;
                k += s - 1;
            } else if (h[k] == -2) {
                continue;
            } else {
                sk_tok[n] = t[k];
                sk_val[n] = (double)(float)v[k];
                n++;
            }
        }
        sk_len[e] = n - sk_off[e];
    }
    return n;
}

/* cl_gpu_eval.cu:165-186 (skeleton interpreter; OP_REF loads a column). */
void sro_predict32_cached(const int *tok, const double *val, int len, const float *X, int nv,
                          int64_t N, const float *cache, float *out)
{
    float stk[SRO_MAX_STACK];
    for (int64_t dp = 0; dp < N; dp++) {
        int sp = 0;
        for (int i = len - 1; i >= 0; i--) {
            int t = tok[i];
            if (t == SRO_OP_REF) {
                stk[sp++] = cache[(int64_t)(int)val[i] * N + dp];
            } else if (t == SRO_TOK_CONST) {
                stk[sp++] = (float)val[i];
            } else if (t == SRO_TOK_VAR) {
                int v = (int)val[i];
                stk[sp++] = v < nv ? X[(int64_t)v * N + dp] : 0.0f;
            } else if (t > 0) {
                float a = stk[--sp], b = 0.0f;
                if (sro_is_binary(t)) b = stk[--sp];
                stk[sp++] = sro_op32(t, a, b);
            }
        }
        out[dp] = stk[0];
    }
}

/* ------------------------------------------------------------------------ */
/* CPU subtree evaluator, FP64 (cpu_subtree.cpp:83-501)                       */
/* ------------------------------------------------------------------------ */

#define SUB_STACK 1100 /* > MAX_EXPR_LEN / 2 + 1: the deepest right-to-left stack */

/* eval_tree_cpu (cpu_subtree.cpp:83-132): right-to-left scan; with hints, -2
 * skips a token inside a reused subtree and h >= 0 pushes column h at dp.
 * x[i] = X[i*N + dp] for i <= nv (the reference's per-point copy). A stack
 * underflow (malformed stream) gives NaN instead of reading outside. */
static double sub_eval_point(const int *tok, const double *val, const int *hints, int len, const double *x,
                             int nx, double *const *cols, int64_t dp)
{
    double st[SUB_STACK];
    int sp = 0;
    for (int i = len - 1; i >= 0; i--) {
        if (hints) {
            if (hints[i] == -2) continue;
            if (hints[i] >= 0) {
                st[sp++] = cols[hints[i]][dp];
                continue;
            }
        }
        const int t = tok[i];
        if (t > 0) {
            const int bin = t < 10;
            if (sp < 1 + bin) return NAN;
            double a = st[--sp], b = 0.0;
            if (bin) b = st[--sp];
            st[sp++] = sro_op64(t, a, b);
        } else if (t == SRO_TOK_CONST) {
            st[sp++] = val[i];
        } else if (t == SRO_TOK_VAR) {
            const int v = (int)val[i];
            st[sp++] = (v >= 0 && v < nx) ? x[v] : NAN;
        }
        if (sp >= SUB_STACK) return NAN;
    }
    return sp > 0 ? st[0] : NAN;
}

typedef struct {
    int phase; /* 0: fill columns, datapoints round robin; 1: expressions round robin */
    int worker, threads;
    /* columns to fill: n_fill subtrees (tokens, values, sizes), written to fill_col */
    int n_fill;
    const int *const *fill_tok;
    const double *const *fill_val;
    const int *fill_len;
    double *const *fill_col;
    /* expressions */
    const int *tok;
    const double *val;
    const int *off;
    const int *len;
    const int *hints; /* packed like tok, or NULL */
    int E;
    double *const *cols;
    const double *X;
    int nv;
    int64_t N;
    double *pred;
} sub_job;

static void *sub_worker(void *arg)
{
    sub_job *j = (sub_job *)arg;
    double x[64];
    const int nx = j->nv + 1 < 64 ? j->nv + 1 : 64;
    if (j->phase == 0) {
        /* cpu_subtree.cpp:191-215: every new subtree's column, datapoints
         * dealt round robin to the workers */
        for (int64_t dp = j->worker; dp < j->N; dp += j->threads) {
            for (int i = 0; i < nx; i++) x[i] = j->X[(int64_t)i * j->N + dp];
            for (int s = 0; s < j->n_fill; s++)
                j->fill_col[s][dp] = sub_eval_point(j->fill_tok[s], j->fill_val[s], NULL, j->fill_len[s], x, nx,
                                                    NULL, 0);
        }
    } else {
        /* cpu_subtree.cpp:241-270: expressions dealt round robin */
        for (int e = j->worker; e < j->E; e += j->threads) {
            const int *t = j->tok + j->off[e];
            const double *v = j->val + j->off[e];
            const int *h = j->hints ? j->hints + j->off[e] : NULL;
            double *out = j->pred + (int64_t)e * j->N;
            for (int64_t dp = 0; dp < j->N; dp++) {
                for (int i = 0; i < nx; i++) x[i] = j->X[(int64_t)i * j->N + dp];
                out[dp] = sub_eval_point(t, v, h, j->len[e], x, nx, j->cols, dp);
            }
        }
    }
    return NULL;
}

static void sub_run(sub_job proto, int threads)
{
    if (threads < 1) threads = 1;
    if (threads > 64) threads = 64;
    pthread_t th[64];
    sub_job jobs[64];
    for (int w = 0; w < threads; w++) {
        jobs[w] = proto;
        jobs[w].worker = w;
        jobs[w].threads = threads;
        pthread_create(&th[w], NULL, sub_worker, &jobs[w]);
    }
    for (int w = 0; w < threads; w++) pthread_join(th[w], NULL);
}

int sro_subtree_predict64(const int *tok, const double *val, const int *off, const int *len, int E,
                          const double *X, int nv, int64_t N, int threads, double *pred, int *hints_out)
{
    /* cpu_subtree.cpp:152-167: min_size 3, min_freq 5 (2 below 10 expressions) */
    sro_detect_result det;
    sro_detect(tok, val, off, len, E, 3, E < 10 ? 2 : 5, &det);
    const int S = det.num_subs;
    double **cols = NULL;
    const int **ft = NULL;
    const double **fv = NULL;
    if (S > 0) {
        cols = (double **)malloc(sizeof(double *) * S);
        ft = (const int **)malloc(sizeof(int *) * S);
        fv = (const double **)malloc(sizeof(double *) * S);
        for (int s = 0; s < S; s++) {
            cols[s] = (double *)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1));
            ft[s] = tok + off[det.sub_expr[s]] + det.sub_start[s];
            fv[s] = val + off[det.sub_expr[s]] + det.sub_start[s];
        }
        sub_job fill = {0};
        fill.phase = 0;
        fill.n_fill = S;
        fill.fill_tok = ft;
        fill.fill_val = fv;
        fill.fill_len = det.sub_size;
        fill.fill_col = cols;
        fill.X = X;
        fill.nv = nv;
        fill.N = N;
        sub_run(fill, threads);
    }
    sub_job ev = {0};
    ev.phase = 1;
    ev.tok = tok;
    ev.val = val;
    ev.off = off;
    ev.len = len;
    ev.hints = S > 0 ? det.hints : NULL; /* cpu_subtree.cpp:262-265 */
    ev.E = E;
    ev.cols = cols;
    ev.X = X;
    ev.nv = nv;
    ev.N = N;
    ev.pred = pred;
    sub_run(ev, threads);
    if (hints_out) {
        int64_t T = 0;
        for (int e = 0; e < E; e++) T = off[e] + len[e] > T ? off[e] + len[e] : T;
        memcpy(hints_out, det.hints, sizeof(int) * (size_t)T);
    }
    for (int s = 0; s < S; s++) free(cols[s]);
    free(cols);
    free(ft);
    free(fv);
    sro_detect_free(&det);
    return S;
}

/* ---- stateful form: a cache of subtree columns kept across calls ---------- */

struct sro_subtree_state {
    int n, cap;
    uint64_t *key;
    int *size;
    int **tok;
    double **val;
    double **col;
    int64_t N;
    int *slot; /* open-addressing index over key, 2*cap entries, -1 = empty */
    int nslot;
};

sro_subtree_state *sro_subtree_state_create(void)
{
    return (sro_subtree_state *)calloc(1, sizeof(sro_subtree_state));
}

static void state_clear(sro_subtree_state *s)
{
    for (int i = 0; i < s->n; i++) {
        free(s->tok[i]);
        free(s->val[i]);
        free(s->col[i]);
    }
    free(s->key);
    free(s->size);
    free(s->tok);
    free(s->val);
    free(s->col);
    free(s->slot);
    memset(s, 0, sizeof(*s));
}

void sro_subtree_state_free(sro_subtree_state *s)
{
    if (!s) return;
    state_clear(s);
    free(s);
}

int sro_subtree_state_size(const sro_subtree_state *s) { return s ? s->n : 0; }

static int same_body(const int *ta, const double *va, const int *tb, const double *vb, int size)
{
    return memcmp(ta, tb, sizeof(int) * size) == 0 && memcmp(va, vb, sizeof(double) * size) == 0;
}

static int state_find(const sro_subtree_state *s, uint64_t h, const int *t, const double *v, int size)
{
    if (!s->nslot) return -1;
    for (uint64_t i = h & (uint64_t)(s->nslot - 1);; i = (i + 1) & (uint64_t)(s->nslot - 1)) {
        const int id = s->slot[i];
        if (id < 0) return -1;
        if (s->key[id] == h && s->size[id] == size && same_body(s->tok[id], s->val[id], t, v, size)) return id;
    }
}

static void state_index(sro_subtree_state *s, int id)
{
    for (uint64_t i = s->key[id] & (uint64_t)(s->nslot - 1);; i = (i + 1) & (uint64_t)(s->nslot - 1))
        if (s->slot[i] < 0) {
            s->slot[i] = id;
            return;
        }
}

static int state_add(sro_subtree_state *s, uint64_t h, const int *t, const double *v, int size)
{
    if (s->n == s->cap) {
        const int cap = s->cap ? 2 * s->cap : 256;
        s->key = (uint64_t *)realloc(s->key, sizeof(uint64_t) * cap);
        s->size = (int *)realloc(s->size, sizeof(int) * cap);
        s->tok = (int **)realloc(s->tok, sizeof(int *) * cap);
        s->val = (double **)realloc(s->val, sizeof(double *) * cap);
        s->col = (double **)realloc(s->col, sizeof(double *) * cap);
        s->cap = cap;
        free(s->slot);
        s->nslot = 4 * cap;
        s->slot = (int *)malloc(sizeof(int) * s->nslot);
        for (int i = 0; i < s->nslot; i++) s->slot[i] = -1;
        for (int i = 0; i < s->n; i++) state_index(s, i);
    }
    const int id = s->n++;
    s->key[id] = h;
    s->size[id] = size;
    s->tok[id] = (int *)malloc(sizeof(int) * size);
    s->val[id] = (double *)malloc(sizeof(double) * size);
    memcpy(s->tok[id], t, sizeof(int) * size);
    memcpy(s->val[id], v, sizeof(double) * size);
    s->col[id] = NULL;
    state_index(s, id);
    return id;
}

typedef struct {
    uint64_t h;
    int e, j, size;
    int cls; /* equal-body class within its hash run */
} sub_cand;

static int sub_cand_cmp(const void *a, const void *b)
{
    const sub_cand *x = (const sub_cand *)a, *y = (const sub_cand *)b;
    if (x->h != y->h) return x->h < y->h ? -1 : 1;
    if (x->cls != y->cls) return x->cls < y->cls ? -1 : 1;
    if (x->e != y->e) return x->e < y->e ? -1 : 1;
    return (x->j > y->j) - (x->j < y->j);
}

typedef struct {
    int first, count, size, id;
    uint64_t h;
} sub_group;

static int sub_group_cmp(const void *a, const void *b)
{
    const sub_group *x = (const sub_group *)a, *y = (const sub_group *)b;
    if (x->size != y->size) return x->size > y->size ? -1 : 1;
    return (x->h > y->h) - (x->h < y->h);
}

int sro_subtree_state_predict64(sro_subtree_state *s, const int *tok, const double *val, const int *off,
                                const int *len, int E, const double *X, int nv, int64_t N, int threads,
                                double *pred, int *hints_out)
{
    if (s->N != N) state_clear(s), s->N = N; /* a cache belongs to one dataset */
    /* cpu_subtree.cpp:462-476 -> detect.cpp:391-560 with min_size 2 and
     * min_freq 5 (2 below 10 expressions); candidates grouped by the flat
     * hash of the subtree itself and then by equal tokens and values (the
     * flat hash collides on high-bit value differences). */
    const int min_size = 2, min_freq = E < 10 ? 2 : 5;
    int64_t T = 0;
    for (int e = 0; e < E; e++) T = off[e] + len[e] > T ? off[e] + len[e] : T;
    int *hints = hints_out ? hints_out : (int *)malloc(sizeof(int) * (size_t)(T > 0 ? T : 1));
    for (int e = 0; e < E; e++)
        for (int j = 0; j < len[e]; j++) hints[off[e] + j] = -1;
    int64_t nc = 0, capc = 1024;
    sub_cand *cand = (sub_cand *)malloc(sizeof(sub_cand) * capc);
    for (int e = 0; e < E; e++) {
        const int *t = tok + off[e];
        const double *v = val + off[e];
        for (int j = 0; j < len[e]; j++) {
            const int size = sro_subtree_size(t, j, len[e]);
            if (size < min_size) continue;
            if (nc == capc) cand = (sub_cand *)realloc(cand, sizeof(sub_cand) * (capc *= 2));
            cand[nc].h = sro_hash_flat(t + j, v + j, size);
            cand[nc].e = e;
            cand[nc].j = j;
            cand[nc].size = size;
            cand[nc].cls = 0;
            nc++;
        }
    }
    qsort(cand, (size_t)nc, sizeof(sub_cand), sub_cand_cmp);
    /* split each hash run into equal-body classes, numbered by first occurrence */
    for (int64_t a = 0; a < nc;) {
        int64_t b = a + 1;
        while (b < nc && cand[b].h == cand[a].h) b++;
        int ncls = 0;
        for (int64_t c = a; c < b; c++) {
            const int *tc = tok + off[cand[c].e] + cand[c].j;
            const double *vc = val + off[cand[c].e] + cand[c].j;
            int k = -1;
            for (int64_t r = a; r < c && k < 0; r++) {
                if (cand[r].size == cand[c].size &&
                    same_body(tok + off[cand[r].e] + cand[r].j, val + off[cand[r].e] + cand[r].j, tc, vc,
                              cand[c].size))
                    k = cand[r].cls;
            }
            cand[c].cls = k >= 0 ? k : ncls++;
        }
        if (ncls > 1) qsort(cand + a, (size_t)(b - a), sizeof(sub_cand), sub_cand_cmp);
        a = b;
    }
    /* groups in ascending hash order (a std::map's order): cached ones are
     * hits, new ones of frequency >= min_freq join the cache in that order */
    int ng = 0, capg = 256, old_n = s->n;
    sub_group *g = (sub_group *)malloc(sizeof(sub_group) * capg);
    for (int64_t a = 0; a < nc;) {
        int64_t b = a + 1;
        while (b < nc && cand[b].h == cand[a].h && cand[b].cls == cand[a].cls) b++;
        const int *ta = tok + off[cand[a].e] + cand[a].j;
        const double *va = val + off[cand[a].e] + cand[a].j;
        int id = state_find(s, cand[a].h, ta, va, cand[a].size);
        if (id < 0 && b - a >= min_freq)
            id = state_add(s, cand[a].h, tok + off[cand[a].e] + cand[a].j, val + off[cand[a].e] + cand[a].j,
                           cand[a].size);
        if (id >= 0) {
            if (ng == capg) g = (sub_group *)realloc(g, sizeof(sub_group) * (capg *= 2));
            g[ng].first = (int)a;
            g[ng].count = (int)(b - a);
            g[ng].size = cand[a].size;
            g[ng].id = id;
            g[ng].h = cand[a].h;
            ng++;
        }
        a = b;
    }
    /* larger subtrees first; an occurrence overlapping a marked one is left
     * to the larger subtree (detect.cpp:522-555) */
    qsort(g, (size_t)ng, sizeof(sub_group), sub_group_cmp);
    for (int k = 0; k < ng; k++)
        for (int c = g[k].first; c < g[k].first + g[k].count; c++) {
            int *h = hints + off[cand[c].e] + cand[c].j;
            int overlap = 0;
            for (int q = 0; q < cand[c].size && !overlap; q++) overlap = h[q] != -1;
            if (overlap) continue;
            h[0] = g[k].id;
            for (int q = 1; q < cand[c].size; q++) h[q] = -2;
        }
    free(g);
    free(cand);
    /* cpu_subtree.cpp:478-495: columns of the new entries, then every expression */
    const int nn = s->n - old_n;
    if (nn > 0) {
        for (int i = old_n; i < s->n; i++) s->col[i] = (double *)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1));
        sub_job fill = {0};
        fill.phase = 0;
        fill.n_fill = nn;
        fill.fill_tok = (const int *const *)(s->tok + old_n);
        fill.fill_val = (const double *const *)(s->val + old_n);
        fill.fill_len = s->size + old_n;
        fill.fill_col = s->col + old_n;
        fill.X = X;
        fill.nv = nv;
        fill.N = N;
        sub_run(fill, threads);
    }
    sub_job ev = {0};
    ev.phase = 1;
    ev.tok = tok;
    ev.val = val;
    ev.off = off;
    ev.len = len;
    ev.hints = s->n > 0 ? hints : NULL;
    ev.E = E;
    ev.cols = s->col;
    ev.X = X;
    ev.nv = nv;
    ev.N = N;
    ev.pred = pred;
    sub_run(ev, threads);
    if (!hints_out) free(hints);
    return nn;
}

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), written from the paper  */
/* ------------------------------------------------------------------------ */

void sro_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void sro_philox_stream_init(sro_philox_stream *s, uint64_t seed, uint32_t stream,
                            uint32_t gen, uint32_t domain)
{
    s->key[0] = (uint32_t)seed;
    s->key[1] = (uint32_t)(seed >> 32);
    s->c1 = stream;
    s->c2 = gen;
    s->c3 = domain;
    s->word = 0;
}

uint32_t sro_philox_next(sro_philox_stream *s)
{
    uint32_t ctr[4] = {s->word >> 2, s->c1, s->c2, s->c3};
    uint32_t out[4];
    sro_philox4x32_10(ctr, s->key, out);
    uint32_t w = out[s->word & 3];
    s->word++;
    return w;
}

/* uniform in [lo, hi): lo + (hi - lo) * u, u = top 24 bits / 2^24.  Each
 * product and sum rounds once (no contraction), identically on the GPU. */
static float philox_uniform(sro_rng *r, float lo, float hi)
{
    uint32_t w = sro_philox_next((sro_philox_stream *)r->state);
    float u = (float)(w >> 8) * (1.0f / 16777216.0f);
    float span = hi - lo;
    float t = span * u;
    return lo + t;
}

/* inclusive [lo, hi] by multiply-shift; no draw when hi <= lo (utils.cpp:973-979). */
static int philox_randint(sro_rng *r, int lo, int hi)
{
    if (hi <= lo) return lo;
    uint32_t w = sro_philox_next((sro_philox_stream *)r->state);
    uint32_t range = (uint32_t)(hi - lo) + 1u;
    return lo + (int)(((uint64_t)w * range) >> 32);
}

void sro_rng_bind_philox(sro_rng *r, sro_philox_stream *s)
{
    r->uniform = philox_uniform;
    r->randint = philox_randint;
    r->state = s;
}

/* ------------------------------------------------------------------------ */
/* GA operators (utils.cpp:934-1272)                                         */
/* ------------------------------------------------------------------------ */

void sro_ga_params_default(sro_ga_params *p)
{
    static const int un[18] = {10, 11, 12, 13, 14, 15, 16, 17, 18,
                               19, 20, 21, 22, 23, 24, 25, 26, 27};
    static const int bi[9] = {1, 2, 3, 4, 5, 6, 7, 8, 9};
    memset(p, 0, sizeof *p);
    p->num_vars = 1;
    p->max_tokens = 64;
    p->max_depth = 5;
    p->const_min = -5.0f;
    p->const_max = 5.0f;
    p->prob_leaf = 0.1f;
    p->prob_const_leaf = 0.5f;
    p->prob_unary = 0.35f;
    p->n_unary = 18;
    p->n_binary = 9;
    memcpy(p->unary_ops, un, sizeof un);
    memcpy(p->binary_ops, bi, sizeof bi);
}

/* opcodes.h:42-78: unknown opcodes count as binary; opcode 29 is unary. */
int sro_op_arity(int op)
{
    if (op >= 10 && op <= 27) return 1;
    if (op == SRO_OP_EXP_CLAMPED) return 1;
    return 2;
}

typedef struct {
    const sro_ga_params *p;
    int max_tokens;
    int left;
    int num_vars;
    int *tok;
    float *val;
    int n;
    sro_rng *r;
} build_ctx;

/* utils.cpp:995-1016 */
static void emit_leaf(build_ctx *c)
{
    if (c->left <= 0) return;
    int choose_const = c->r->uniform(c->r, 0.0f, 1.0f) < c->p->prob_const_leaf;
    if (c->num_vars <= 0) choose_const = 1;
    if (choose_const) {
        c->tok[c->n] = SRO_TOK_CONST;
        c->val[c->n] = c->r->uniform(c->r, c->p->const_min, c->p->const_max);
    } else {
        int v = c->r->randint(c->r, 0, c->num_vars - 1);
        c->tok[c->n] = SRO_TOK_VAR;
        c->val[c->n] = (float)v;
    }
    c->n++;
    c->left -= 1;
}

/* utils.cpp:1018-1084 */
static void build_node(build_ctx *c, int depth)
{
    if (c->left <= 0) return;
    int must_leaf = depth >= c->p->max_depth || c->left == 1;
    if (!must_leaf) {
        if (c->r->uniform(c->r, 0.0f, 1.0f) < c->p->prob_leaf) must_leaf = 1;
    }
    if (must_leaf) {
        emit_leaf(c);
        return;
    }
    int allow_unary = c->left >= 2;
    int allow_binary = c->left >= 3;
    if (!allow_unary && !allow_binary) {
        emit_leaf(c);
        return;
    }
    float thresh = c->p->prob_unary;
    if (!allow_unary) thresh = 0.0f;
    if (!allow_binary) thresh = 1.0f;
    int pick_unary = c->r->uniform(c->r, 0.0f, 1.0f) < thresh;
    int op, arity;
    if (pick_unary) {
        op = c->p->unary_ops[c->r->randint(c->r, 0, c->p->n_unary - 1)];
        arity = 1;
    } else {
        op = c->p->binary_ops[c->r->randint(c->r, 0, c->p->n_binary - 1)];
        arity = 2;
    }
    c->tok[c->n] = op;
    c->val[c->n] = 0.0f;
    c->n++;
    c->left -= 1;
    for (int i = 0; i < arity; i++) {
        if (c->left <= 0) {
            c->tok[c->n] = SRO_TOK_VAR; /* budget exhausted: fallback leaf */
            c->val[c->n] = 0.0f;
            c->n++;
            continue;
        }
        build_node(c, depth + 1);
    }
}

/* generate_with_limit, utils.cpp:1086-1108 */
int sro_ga_generate(const sro_ga_params *p, int max_tokens, sro_rng *r, int *tok, float *val)
{
    if (max_tokens <= 0) return 0;
    build_ctx c;
    c.p = p;
    c.max_tokens = p->max_tokens < max_tokens ? p->max_tokens : max_tokens;
    c.left = c.max_tokens;
    c.num_vars = p->num_vars > 1 ? p->num_vars : 1;
    c.tok = tok;
    c.val = val;
    c.n = 0;
    c.r = r;
    build_node(&c, 0);
    if (c.n == 0) {
        tok[0] = SRO_TOK_CONST;
        val[0] = 0.0f;
        c.n = 1;
    }
    return c.n;
}

/* utils.cpp:1110-1129 (op_arity based; terminals are CONST and VAR only). */
int sro_ga_subtree_size(const int *tok, int len, int start)
{
    if (start < 0 || start >= len) return 0;
    int t = tok[start];
    int ar = (t == SRO_TOK_CONST || t == SRO_TOK_VAR) ? 0 : sro_op_arity(t);
    int consumed = 1, cur = start + 1;
    for (int i = 0; i < ar; i++) {
        if (cur >= len) break;
        int ch = sro_ga_subtree_size(tok, len, cur);
        if (ch <= 0) break;
        consumed += ch;
        cur += ch;
    }
    return consumed;
}

/* splice_subtree, utils.cpp:1138-1153 */
static int splice(const int *st, const float *sv, int slen, int begin, int count,
                  const int *it, const float *iv, int ilen, int *dt, float *dv)
{
    int n = 0;
    for (int k = 0; k < begin; k++, n++) { dt[n] = st[k]; dv[n] = sv[k]; }
    for (int k = 0; k < ilen; k++, n++) { dt[n] = it[k]; dv[n] = iv[k]; }
    for (int k = begin + count; k < slen; k++, n++) { dt[n] = st[k]; dv[n] = sv[k]; }
    return n;
}

static int copy_expr(const int *st, const float *sv, int len, int *dt, float *dv)
{
    memcpy(dt, st, sizeof(int) * (size_t)len);
    memcpy(dv, sv, sizeof(float) * (size_t)len);
    return len;
}

#define SRO_GA_TMP 4096

/* mutate_expression, utils.cpp:1163-1204 */
int sro_ga_mutate(const sro_ga_params *p, const int *ptok, const float *pval, int plen,
                  sro_rng *r, int *ctok, float *cval)
{
    if (plen <= 0) return sro_ga_generate(p, p->max_tokens, r, ctok, cval);
    int idx = r->randint(r, 0, plen - 1);
    int old = sro_ga_subtree_size(ptok, plen, idx);
    if (old <= 0) old = 1;
    int base = plen - old;
    int max_new = p->max_tokens - base;
    if (max_new <= 0) return copy_expr(ptok, pval, plen, ctok, cval);
    int nt[SRO_GA_TMP];
    float nv[SRO_GA_TMP];
    int nlen = sro_ga_generate(p, max_new, r, nt, nv);
    return splice(ptok, pval, plen, idx, old, nt, nv, nlen, ctok, cval);
}

/* crossover_expressions, utils.cpp:1206-1272 */
int sro_ga_crossover(const sro_ga_params *p, const int *ltok, const float *lval, int llen,
                     const int *rtok, const float *rval, int rlen, sro_rng *r, int *ctok,
                     float *cval)
{
    if (llen <= 0 || rlen <= 0) return copy_expr(ltok, lval, llen, ctok, cval);
    int li = r->randint(r, 0, llen - 1);
    int lsub = sro_ga_subtree_size(ltok, llen, li);
    if (lsub <= 0) lsub = 1;
    int attempts = rlen < 32 ? rlen : 32;
    for (int a = 0; a < attempts; a++) {
        int ri = r->randint(r, 0, rlen - 1);
        int rsub = sro_ga_subtree_size(rtok, rlen, ri);
        if (rsub <= 0) continue;
        if (llen - lsub + rsub > p->max_tokens) continue;
        return splice(ltok, lval, llen, li, lsub, rtok + ri, rval + ri, rsub, ctok, cval);
    }
    return copy_expr(ltok, lval, llen, ctok, cval);
}

static const float *g_fit_for_sort;

static int fit_cmp(const void *a, const void *b)
{
    int i = *(const int *)a, j = *(const int *)b;
    float fi = g_fit_for_sort[i], fj = g_fit_for_sort[j];
    int ni = isnan(fi), nj = isnan(fj);
    if (ni != nj) return ni ? 1 : -1;
    if (!ni && fi != fj) return fi < fj ? -1 : 1;
    return (i > j) - (i < j);
}

/* cl_sr_system.cpp:287-291: ascending fitness, NaN last; ties by index. */
void sro_fitness_order(const float *fit, int n, int *order)
{
    for (int i = 0; i < n; i++) order[i] = i;
    g_fit_for_sort = fit;
    qsort(order, (size_t)n, sizeof(int), fit_cmp);
    g_fit_for_sort = NULL;
}

/* cl_sr_system.cpp:305-313: winner of t draws, strict < keeps the earlier. */
static int tournament(sro_rng *r, const float *sfit, int pop, int t)
{
    int best = r->randint(r, 0, pop - 1);
    for (int k = 0; k < t - 1; k++) {
        int cand = r->randint(r, 0, pop - 1);
        if (sfit[cand] < sfit[best]) best = cand;
    }
    return best;
}

/* Selection + variation loop of cl_sr_system.cpp:285-330 with the random
 * stream of child c supplied by rng_for(ud, c); parent1 / parent2 (nullable)
 * receive the parents of every child (an elite: its index and -2; a mutant:
 * -1 as parent2). */
static int next_generation(const sro_ga_params *p, const sro_select_params *sp, int pop, int cap,
                           const int *tok, const float *val, const int *len, const float *fit,
                           sro_rng *(*rng_for)(void *, int), void *ud, int *out_tok, float *out_val,
                           int *out_len, int *parent1, int *parent2)
{
    int *order = (int *)malloc(sizeof(int) * (size_t)pop);
    float *sfit = (float *)malloc(sizeof(float) * (size_t)pop);
    sro_fitness_order(fit, pop, order);
    for (int k = 0; k < pop; k++) sfit[k] = fit[order[k]];
    int elites = (int)(pop * sp->elite_frac);
    if (elites < 1) elites = 1;
    if (elites > pop) elites = pop;
    int rc = 0;
    for (int c = 0; c < pop; c++) {
        int *ct = out_tok + (size_t)c * cap;
        float *cv = out_val + (size_t)c * cap;
        int n;
        if (c < elites) {
            int s = order[c];
            n = copy_expr(tok + (size_t)s * cap, val + (size_t)s * cap, len[s], ct, cv);
            if (parent1) parent1[c] = s;
            if (parent2) parent2[c] = -2;
        } else {
            sro_rng *r = rng_for(ud, c);
            int b1 = order[tournament(r, sfit, pop, sp->tournament)];
            int b2 = -1;
            int tmp[SRO_GA_TMP];
            float tv[SRO_GA_TMP];
            if (r->uniform(r, 0.0f, 1.0f) < sp->p_mutation) {
                n = sro_ga_mutate(p, tok + (size_t)b1 * cap, val + (size_t)b1 * cap, len[b1], r, tmp, tv);
            } else {
                b2 = order[tournament(r, sfit, pop, sp->tournament)];
                n = sro_ga_crossover(p, tok + (size_t)b1 * cap, val + (size_t)b1 * cap, len[b1],
                                     tok + (size_t)b2 * cap, val + (size_t)b2 * cap, len[b2], r, tmp, tv);
            }
            if (n > cap) {
                rc = -1;
                n = cap;
            }
            memcpy(ct, tmp, sizeof(int) * (size_t)n);
            memcpy(cv, tv, sizeof(float) * (size_t)n);
            if (parent1) parent1[c] = b1;
            if (parent2) parent2[c] = b2;
        }
        out_len[c] = n;
    }
    free(order);
    free(sfit);
    return rc;
}

/* Per-child Philox streams (seed, child, gen, domain 1). */
typedef struct {
    uint64_t seed;
    uint32_t gen;
    sro_philox_stream st;
    sro_rng r;
} philox_children;

static sro_rng *philox_child(void *ud, int c)
{
    philox_children *pc = (philox_children *)ud;
    sro_philox_stream_init(&pc->st, pc->seed, (uint32_t)c, pc->gen, 1u);
    sro_rng_bind_philox(&pc->r, &pc->st);
    return &pc->r;
}

int sro_ga_next_generation(const sro_ga_params *p, const sro_select_params *sp, int pop,
                           int cap, const int *tok, const float *val, const int *len,
                           const float *fit, uint64_t seed, uint32_t gen, int *out_tok,
                           float *out_val, int *out_len)
{
    philox_children pc;
    pc.seed = seed;
    pc.gen = gen;
    return next_generation(p, sp, pop, cap, tok, val, len, fit, philox_child, &pc, out_tok, out_val, out_len,
                           NULL, NULL);
}

int sro_ga_next_generation_hooked(const sro_ga_params *p, const sro_select_params *sp, int pop, int cap,
                                  const int *tok, const float *val, const int *len, const float *fit,
                                  sro_rng *(*rng_for)(void *ud, int child), void *ud, int *out_tok,
                                  float *out_val, int *out_len, int *parent1, int *parent2)
{
    return next_generation(p, sp, pop, cap, tok, val, len, fit, rng_for, ud, out_tok, out_val, out_len, parent1,
                           parent2);
}

int sro_ga_init_population(const sro_ga_params *p, int pop, int cap, uint64_t seed,
                           int *out_tok, float *out_val, int *out_len)
{
    int rc = 0;
    for (int i = 0; i < pop; i++) {
        sro_philox_stream st;
        sro_rng r;
        sro_philox_stream_init(&st, seed, (uint32_t)i, 0xFFFFFFFFu, 2u);
        sro_rng_bind_philox(&r, &st);
        int tmp[SRO_GA_TMP];
        float tv[SRO_GA_TMP];
        int n = sro_ga_generate(p, p->max_tokens, &r, tmp, tv);
        if (n > cap) {
            rc = -1;
            n = cap;
        }
        memcpy(out_tok + (size_t)i * cap, tmp, sizeof(int) * (size_t)n);
        memcpy(out_val + (size_t)i * cap, tv, sizeof(float) * (size_t)n);
        out_len[i] = n;
    }
    return rc;
}

/* ---- evogp layout --------------------------------------------------------- */
void sro_evogp_to_prefix(int pop, int max_len, const float *value, const short *type, const short *size,
                         int cap, int *tok, float *val, int *len)
{
    for (int i = 0; i < pop; ++i) {
        const float *vi = value + (size_t)i * max_len;
        const short *ti = type + (size_t)i * max_len;
        const int n = size[(size_t)i * max_len];
        int *to = tok + (size_t)i * cap;
        float *vo = val + (size_t)i * cap;
        if (n <= 0 || n > max_len || n > cap) {
            len[i] = 0;
            continue;
        }
        for (int t = 0; t < n; ++t) {
            const int kind = ti[t] & 0x7f;
            if (kind == 1) {        /* CONST */
                to[t] = 0;
                vo[t] = vi[t];
            } else if (kind == 0) { /* VAR */
                to[t] = -1;
                vo[t] = vi[t];
            } else {                /* function: the opcode is the value */
                to[t] = (int)vi[t];
                vo[t] = 0.0f;
            }
        }
        len[i] = n;
    }
}

/* End of the subtree starting at t (one past it), or -1 past the stream. */
static int evogp_extent(const int *tok, int n, int t)
{
    int need = 1, k = t;
    while (need > 0) {
        if (k >= n) return -1;
        const int x = tok[k++];
        need += (x <= 0 ? 0 : (x < 10 ? 2 : 1)) - 1;
    }
    return k;
}

void sro_prefix_to_evogp(int pop, int cap, const int *tok, const float *val, const int *len, int max_len,
                         float *value, short *type, short *size)
{
    for (int i = 0; i < pop; ++i) {
        const int *ti = tok + (size_t)i * cap;
        const float *vi = val + (size_t)i * cap;
        float *vo = value + (size_t)i * max_len;
        short *to = type + (size_t)i * max_len;
        short *so = size + (size_t)i * max_len;
        const int n = len[i] < cap ? len[i] : cap;
        int bad = n <= 0 || n > max_len || n > 32767;
        for (int t = 0; t < n && !bad; ++t)
            bad = !(ti[t] == 0 || ti[t] == -1 || (ti[t] >= 1 && ti[t] <= 27) || ti[t] == 29);
        if (bad || evogp_extent(ti, n, 0) != n) {
            so[0] = 0;
            continue;
        }
        for (int t = 0; t < n; ++t) {
            const int x = ti[t];
            to[t] = (short)(x == 0 ? 1 : x == -1 ? 0 : (x < 10 && x > 0) ? 3 : 2);
            vo[t] = (x == 0 || x == -1) ? vi[t] : (float)x;
            so[t] = (short)(evogp_extent(ti, n, t) - t);
        }
    }
}
