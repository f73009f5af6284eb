// sr_format.h -- token format of the reference and the bytecode the batched
// interpreter runs.  Shared by host code and device kernels.
//
// Token format (opcodes.h:10-40): TOK_CONST = 0 (value = constant),
// TOK_VAR = -1 (value = variable index), binary ops 1..9, unary ops 10..27.
// OP_REF = -2 marks a cached-subtree reference in a skeleton
// (cl_gpu_eval.cu:11; value = cache slot).  OP_STORE = -3 (internal to the
// generational cache, never accepted from a caller) is a unary identity whose
// operand value is also written to cache column 'value'.  Opcode 29 is the
// clamped exp (evolution_tests.cpp:40) that the reference evaluator does not
// define.
//
// Bytecode ("register program"): the prefix tree is re-ordered into an
// accumulator program.  Leaves are fused into the instruction that consumes
// them, the larger subtree of a binary node is evaluated first (Sethi-Ullman),
// so the per-lane stack holds only the values that must wait for a sibling.
// Re-ordering the evaluation of independent subtrees does not change any value:
// every node still sees bit-identical operands.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define SR_HD __host__ __device__ inline
#else
#define SR_HD inline
#endif

namespace sr {

constexpr int TOK_CONST = 0;
constexpr int TOK_VAR = -1;
constexpr int OP_REF = -2;
constexpr int OP_STORE = -3;

constexpr int OP_ADD = 1, OP_SUB = 2, OP_MUL = 3, OP_DIV = 4, OP_POW = 5, OP_MIN = 6,
              OP_MAX = 7, OP_LOOSE_DIV = 8, OP_LOOSE_POW = 9;
constexpr int OP_SIN = 10, OP_COS = 11, OP_TAN = 12, OP_SINH = 13, OP_COSH = 14,
              OP_TANH = 15, OP_EXP = 16, OP_LOG = 17, OP_INV = 18, OP_ASIN = 19,
              OP_ACOS = 20, OP_ATAN = 21, OP_LOOSE_LOG = 22, OP_LOOSE_INV = 23,
              OP_ABS = 24, OP_NEG = 25, OP_SQRT = 26, OP_LOOSE_SQRT = 27;
constexpr int OP_EXP_CLAMPED = 29;

constexpr int MAX_EXPR_LEN = 1024;  // longest token stream accepted
constexpr int MAX_VARS = 32;        // utils.h:14 MAX_VAR_NUM

// Interpreter arity (cl_cpu_eval.cpp:82): 1..9 binary, others unary.
SR_HD bool is_binary(int t) { return t >= 1 && t <= 9; }
SR_HD bool is_leaf(int t) { return t == TOK_CONST || t == TOK_VAR || t == OP_REF; }
SR_HD bool is_valid_token(int t)
{
    return is_leaf(t) || (t >= 1 && t <= 27) || t == OP_EXP_CLAMPED;
}
// Bytecode op of the OP_STORE pseudo-token: a unary instruction on the
// accumulator whose arg is the destination column; the whole code word is
// always STORE_CODE (its operand is a subtree, never a fused leaf).
constexpr int OPC_STORE = 31;
// IEEE add/mul are commutative bit for bit; everything else keeps its order.
SR_HD bool is_commutative(int t) { return t == OP_ADD || t == OP_MUL; }

// ---- bytecode --------------------------------------------------------------
// code: bits 0..7 op (0 = LOAD), bits 8..10 source of 'o', bit 11 push top
// before the load, bit 12 swap(top, o) before a binary op.
// arg : const bits / variable index / cache slot.
constexpr int SRC_NONE = 0, SRC_CONST = 1, SRC_VAR = 2, SRC_REF = 3, SRC_POP = 4;
constexpr int BC_PUSH = 1 << 11, BC_SWAP = 1 << 12;
// Bits 13..14: instruction class (0 = leaf load, 1 = binary op, 2 = unary op),
// so the interpreter dispatches operand staging on bits 8..14 in one switch.
constexpr int BC_CLS_SHIFT = 13;
constexpr int CL_LOAD = 0, CL_BIN = 1, CL_UN = 2;

SR_HD int bc_op(int code) { return code & 0xff; }
SR_HD int bc_src(int code) { return (code >> 8) & 7; }
SR_HD int bc_cls(int code) { return (code >> BC_CLS_SHIFT) & 3; }
// Staging form: src | push << 3 | swap << 4 | class << 5.
SR_HD int bc_form(int code) { return (code >> 8) & 0x7f; }
constexpr int bc_form_of(int src, int push, int swap, int cls)
{
    return src | (push ? 8 : 0) | (swap ? 16 : 0) | (cls << 5);
}
constexpr int32_t STORE_CODE = OPC_STORE | (SRC_NONE << 8) | (CL_UN << BC_CLS_SHIFT);
// OP_REF operands prefetched a tile ahead: the first kRefPrefetch OP_REF
// instructions of a cache skeleton carry j + 1 in bits 16..18 (j = their
// order) and the program's slot list names their columns.  Bits 16.. are
// otherwise used only by the adjoint's parametric programs.
constexpr int kRefPrefetch = 4;
constexpr int BC_PF_SHIFT = 16;
SR_HD int bc_pf(int code) { return (code >> BC_PF_SHIFT) & 7; }

struct Ins {
    int32_t code;
    int32_t arg;
};

SR_HD int32_t f32_bits(float f)
{
    union {
        float f;
        int32_t i;
    } u;
    u.f = f;
    return u.i;
}

// Raw token program for the plain interpreter: {token, payload}.
// payload: const bits, variable index, or cache slot.  A variable index >= nv
// evaluates to 0 (cl_cpu_eval.cpp:71-75) and is lowered to CONST 0 here.
SR_HD int lower_token(int t, float v, int nv, Ins *out)
{
    if (!is_valid_token(t)) return -1;
    if (t == TOK_CONST) {
        out->code = TOK_CONST;
        out->arg = f32_bits(v);
    } else if (t == TOK_VAR) {
        int k = (int)v;
        if (k < 0) return -1;
        if (k >= nv) {
            out->code = TOK_CONST;
            out->arg = 0;
        } else {
            out->code = TOK_VAR;
            out->arg = k;
        }
    } else if (t == OP_REF) {
        int s = (int)v;
        if (s < 0) return -1;
        out->code = OP_REF;
        out->arg = s;
    } else {
        out->code = t;
        out->arg = 0;
    }
    return 0;
}

// Stack depth the raw right-to-left interpreter reaches; -1 when malformed.
// OP_STORE is accepted only with store = true (cache skeletons).
SR_HD int raw_stack_depth(const int32_t *tok, int len, bool store = false)
{
    int sp = 0, mx = 0;
    for (int i = len - 1; i >= 0; --i) {
        int t = tok[i];
        if (!is_valid_token(t) && !(store && t == OP_STORE)) return -1;
        if (is_leaf(t)) {
            sp++;
        } else {
            int need = is_binary(t) ? 2 : 1;
            if (sp < need) return -1;
            sp -= need - 1;
        }
        if (sp > mx) mx = sp;
    }
    return (len > 0 && sp == 1) ? mx : -1;
}

// Leaf -> (src, arg).
SR_HD void leaf_src(int t, float v, int nv, int &src, int32_t &arg)
{
    if (t == TOK_CONST) {
        src = SRC_CONST;
        arg = f32_bits(v);
    } else if (t == TOK_VAR) {
        int k = (int)v;
        if (k >= nv) {
            src = SRC_CONST;
            arg = 0;
        } else {
            src = SRC_VAR;
            arg = k;
        }
    } else {
        src = SRC_REF;
        arg = (int)v;
    }
}

// Emission frame of compile_expr (caller-provided scratch, len + 1 frames).
struct CompileFrame {
    int node, stage, live, second, swap;
};

// Compile one validated prefix expression into bytecode.
//   tok/val : token stream (len tokens, already checked by raw_stack_depth)
//   size/need : scratch arrays of len ints; fr: scratch of len + 1 frames
//   leaf_pos : optional [len] output, token position of the leaf each
//              instruction loads (-1 for none / a popped operand)
// Returns the number of instructions (<= len) and sets *stack_need; -1 on a
// malformed stream.
SR_HD int compile_expr(const int32_t *tok, const float *val, int len, int nv, int *size,
                       int *need, CompileFrame *fr, Ins *out, int *stack_need,
                       int *leaf_pos = nullptr)
{
    if (len <= 0 || len > MAX_EXPR_LEN) return -1;
    // Right-to-left pass: subtree sizes and stack need (Sethi-Ullman).
    for (int i = len - 1; i >= 0; --i) {
        int t = tok[i];
        if (is_leaf(t)) {
            size[i] = 1;
            need[i] = 0;
        } else if (!is_binary(t)) {
            int c = i + 1;
            if (c >= len) return -1;
            size[i] = 1 + size[c];
            need[i] = need[c];
        } else {
            int a = i + 1;
            if (a >= len) return -1;
            int b = a + size[a];
            if (b >= len) return -1;
            size[i] = 1 + size[a] + size[b];
            bool la = is_leaf(tok[a]), lb = is_leaf(tok[b]);
            if (la && lb) need[i] = 0;
            else if (la) need[i] = need[b];
            else if (lb) need[i] = need[a];
            else if (need[a] != need[b]) need[i] = need[a] > need[b] ? need[a] : need[b];
            else need[i] = need[a] + 1;
        }
    }
    if (size[0] != len) return -1;
    *stack_need = need[0];

    // Iterative emission.  Frame stages: 0 = enter, 1 = emit after the single
    // evaluated child, 2 = after first of two subtrees (then evaluate second),
    // 3 = emit after the second subtree.
    int fsp = 0, n = 0;
    fr[fsp++] = {0, 0, 0, 0, 0};
    while (fsp > 0) {
        CompileFrame &f = fr[fsp - 1];
        int t = tok[f.node];
        int push = f.live ? BC_PUSH : 0;
        if (f.stage == 0) {
            if (is_leaf(t)) {
                int src;
                int32_t arg;
                leaf_src(t, val[f.node], nv, src, arg);
                if (leaf_pos) leaf_pos[n] = f.node;
                out[n++] = {0 | (src << 8) | push | (CL_LOAD << BC_CLS_SHIFT), arg};
                --fsp;
                continue;
            }
            if (!is_binary(t)) {
                int c = f.node + 1;
                if (t == OP_STORE && is_leaf(tok[c])) return -1;
                if (is_leaf(tok[c])) {
                    int src;
                    int32_t arg;
                    leaf_src(tok[c], val[c], nv, src, arg);
                    if (leaf_pos) leaf_pos[n] = c;
                    out[n++] = {t | (src << 8) | push | (CL_UN << BC_CLS_SHIFT), arg};
                    --fsp;
                } else {
                    f.stage = 1;
                    f.swap = 0;
                    f.second = -1;  // unary: emit op on top
                    fr[fsp++] = {c, 0, f.live, 0, 0};
                }
                continue;
            }
            int a = f.node + 1, b = a + size[a];
            bool la = is_leaf(tok[a]), lb = is_leaf(tok[b]);
            bool comm = is_commutative(t);
            if (la && lb) {
                int src;
                int32_t arg;
                leaf_src(tok[a], val[a], nv, src, arg);
                if (leaf_pos) leaf_pos[n] = a;
                out[n++] = {0 | (src << 8) | push | (CL_LOAD << BC_CLS_SHIFT), arg};
                leaf_src(tok[b], val[b], nv, src, arg);
                if (leaf_pos) leaf_pos[n] = b;
                out[n++] = {t | (src << 8) | (CL_BIN << BC_CLS_SHIFT), arg};
                --fsp;
            } else if (la) {
                // evaluate B, then top = f(A, top)
                f.stage = 1;
                f.second = a;
                f.swap = comm ? 0 : BC_SWAP;
                fr[fsp++] = {b, 0, f.live, 0, 0};
            } else if (lb) {
                f.stage = 1;
                f.second = b;
                f.swap = 0;
                fr[fsp++] = {a, 0, f.live, 0, 0};
            } else {
                bool a_first = need[a] > need[b];
                f.stage = 2;
                f.second = a_first ? b : a;
                f.swap = (a_first && !comm) ? BC_SWAP : 0;
                fr[fsp++] = {a_first ? a : b, 0, f.live, 0, 0};
            }
            continue;
        }
        if (f.stage == 1) {
            if (f.second < 0) {
                if (leaf_pos) leaf_pos[n] = -1;
                if (t == OP_STORE) out[n++] = {STORE_CODE, (int32_t)val[f.node]};
                else out[n++] = {t | (SRC_NONE << 8) | (CL_UN << BC_CLS_SHIFT), 0};
            } else {
                int src;
                int32_t arg;
                leaf_src(tok[f.second], val[f.second], nv, src, arg);
                if (leaf_pos) leaf_pos[n] = f.second;
                out[n++] = {t | (src << 8) | f.swap |
                                ((is_binary(t) ? CL_BIN : CL_UN) << BC_CLS_SHIFT),
                            arg};
            }
            --fsp;
            continue;
        }
        if (f.stage == 2) {
            f.stage = 3;
            fr[fsp++] = {f.second, 0, 1, 0, 0};
            continue;
        }
        // stage 3: o = pop (first subtree's value), top = second subtree
        if (leaf_pos) leaf_pos[n] = -1;
        out[n++] = {t | (SRC_POP << 8) | f.swap | (CL_BIN << BC_CLS_SHIFT), 0};
        --fsp;
    }
    return n;
}

// ---- parametric programs (constant fitting) ---------------------------------
// Bits 16..19 of a SRC_CONST instruction hold 1 + the free-constant slot it
// reads (0 = a fixed constant); a SRC_POP instruction's arg holds the index of
// the instruction whose output it pops (the one before the matching push),
// which is where the adjoint pass finds that operand on its tape.
constexpr int MAX_SLOTS = 8;
SR_HD int bc_slot(int code) { return ((code >> 16) & 15) - 1; }

// const_slot: per token of the expression, -1 or its free-constant slot.
// Returns 0, or -1 if a slot is out of range.
SR_HD int param_postpass(Ins *ins, int n, const int *leaf_pos, const int32_t *const_slot,
                         int *pop_stack)
{
    int sp = 0;
    for (int i = 0; i < n; ++i) {
        int code = ins[i].code;
        if (code & BC_PUSH) pop_stack[sp++] = i - 1;
        const int src = bc_src(code);
        if (src == SRC_POP) {
            ins[i].arg = pop_stack[--sp];
        } else if (src == SRC_CONST && leaf_pos && const_slot && leaf_pos[i] >= 0) {
            const int s = const_slot[leaf_pos[i]];
            if (s >= MAX_SLOTS) return -1;
            if (s >= 0) ins[i].code = code | ((s + 1) << 16);
        }
    }
    return 0;
}

}  // namespace sr
