// sr_compile_wave.h -- the bytecode compiler of sr_format.h (compile_expr)
// restated for one 64-lane wave: no step is serial over the tokens, so a
// program of n tokens compiles in a few dozen dependent LDS round trips
// instead of several per token.  The device cache planner compiles every
// skeleton with it; the host runs the same code with the lanes taken one
// after the other (the CPU tests compare it with compile_expr).
//
// The algorithm (every per-node step runs on all nodes at once):
//   1. S[k] = sum over tokens before k of (1 - arity); the stream is valid iff
//      every token is valid, S[n] = 1 and S[k] <= 0 for 0 < k < n.
//   2. subtree size: the end of node i is the first j > i with
//      S[j] = S[i] + 1 (S rises by at most 1 per token, so the search skips).
//   3. stack need (Sethi-Ullman, as compile_expr): rounds over the nodes
//      until every node whose children are known is known; rounds <= height.
//   4. instruction count of a subtree = sum over its span of w (1 per
//      operator, 2 for a binary operator of two leaves), a prefix sum W;
//      the children evaluated first / second follow compile_expr's rule, the
//      second one starts after the first one's instructions.
//   5. start of every subtree's instructions = sum of those offsets along
//      the path to the root, by pointer jumping (log2(height) rounds).
//   6. each operator writes its instruction at start + count - 1 (a binary
//      operator of two leaves also its LOAD at start); the first instruction
//      of every second-evaluated subtree takes BC_PUSH.
// The bytecode equals compile_expr's, instruction for instruction.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#endif

#include "sr_format.h"

namespace sr {

constexpr int kWaveLanes = 64;

// LDS (or host) scratch for programs of up to `cap` tokens.  Per-node values
// are 16-bit (|S|, sizes, counts, starts and indexes are <= MAX_EXPR_LEN), so
// more programs fit in LDS at once.
struct WaveCompileScratch {
    int32_t *part;                     // [kWaveLanes]
    int32_t *flag;                     // [4]
    uint32_t *bits;                    // [(cap + 31) / 32] OP_REF operands by position
    int16_t *S;                        // [cap + 1]
    int16_t *size, *need;              // [cap]
    int16_t *W;                        // [cap + 1]
    int16_t *acc0, *acc1, *up0, *up1;  // [cap]
};
static_assert(MAX_EXPR_LEN < 32768, "16-bit compile scratch");

// Scratch ints needed for programs of up to cap tokens.
SR_HD int wave_compile_scratch_ints(int cap) { return kWaveLanes + 4 + (cap + 31) / 32 + (8 * cap + 2 + 1) / 2; }

SR_HD WaveCompileScratch wave_compile_scratch(int32_t *base, int cap)
{
    WaveCompileScratch s;
    s.part = base;
    s.flag = s.part + kWaveLanes;
    s.bits = (uint32_t *)(s.flag + 4);
    s.S = (int16_t *)(s.bits + (cap + 31) / 32);
    s.size = s.S + cap + 1;
    s.need = s.size + cap;
    s.W = s.need + cap;
    s.acc0 = s.W + cap + 1;
    s.acc1 = s.acc0 + cap;
    s.up0 = s.acc1 + cap;
    s.up1 = s.up0 + cap;
    return s;
}

// Lanes of one wave.  each(f) runs f(lane) on every lane, then makes every
// lane's writes visible to all (on the device a workgroup barrier; on the
// host the lanes simply run in order).  State carried from one step to the
// next lives in the scratch arrays; values read outside each() are uniform.
#if defined(__HIPCC__)
struct DeviceWave {
    template <class F>
    __device__ void each(F &&f)
    {
        f((int)threadIdx.x);
        __syncthreads();
    }
};
#endif
struct HostWave {
    template <class F>
    void each(F &&f)
    {
        for (int l = 0; l < kWaveLanes; ++l) f(l);
    }
};

SR_HD int wave_block_lo(int n, int lane)
{
    const int per = (n + kWaveLanes - 1) / kWaveLanes;
    return lane * per < n ? lane * per : n;
}
SR_HD int wave_block_hi(int n, int lane)
{
    const int per = (n + kWaveLanes - 1) / kWaveLanes;
    const int hi = (lane + 1) * per;
    return hi < n ? hi : n;
}

SR_HD int tok_delta(int t) { return is_leaf(t) ? 1 : (is_binary(t) ? -1 : 0); }

// OR into a scratch word shared by the lanes.
SR_HD void wave_or(uint32_t *p, uint32_t v)
{
#if defined(__HIP_DEVICE_COMPILE__)
    atomicOr(p, v);
#else
    *p |= v;
#endif
}

// Number of set bits below position pos, stopping once it reaches cap_count.
SR_HD int bits_below(const uint32_t *bits, int pos, int cap_count)
{
    int r = 0;
    for (int w = 0; w < (pos >> 5) && r < cap_count; ++w) r += __builtin_popcount(bits[w]);
    if (pos & 31) r += __builtin_popcount(bits[pos >> 5] & ((1u << (pos & 31)) - 1u));
    return r;
}

// Compile tok/val[0..n) (OP_REF and OP_STORE accepted); out: [n] (written
// only when the stream is valid).  Returns the instruction count (uniform),
// or -1 for a malformed stream, and sets *stack_need.  Same result as
// raw_stack_depth(tok, n, true) >= 0 followed by compile_expr.  With pf
// (kRefPrefetch entries), the first kRefPrefetch OP_REF operands in program
// order carry j + 1 in bits BC_PF_SHIFT.. and pf[j] = their slot (-1 past
// the last one), as the skeleton launches expect.
template <class Wave>
SR_HD int compile_wave(Wave &w, const int32_t *tok, const float *val, int n, int nv, WaveCompileScratch s,
                       Ins *out, int *stack_need, int32_t *pf = nullptr)
{
    if (n <= 0 || n > MAX_EXPR_LEN) return -1;
    int16_t *S = s.S, *size = s.size, *need = s.need, *W = s.W;
    int32_t *part = s.part, *flag = s.flag;
    // 1. validity and S
    w.each([&](int lane) {
        if (lane < 4) flag[lane] = 0;
        int d = 0;
        for (int k = wave_block_lo(n, lane); k < wave_block_hi(n, lane); ++k) d += tok_delta(tok[k]);
        part[lane] = d;
    });
    w.each([&](int lane) {
        int x = 0;
        for (int l = 0; l < lane; ++l) x += part[l];
        bool bad = false;
        for (int k = wave_block_lo(n, lane); k < wave_block_hi(n, lane); ++k) {
            const int t = tok[k];
            S[k] = x;
            bad = bad || (k > 0 && x > 0) || !(is_valid_token(t) || t == OP_STORE);
            x += tok_delta(t);
        }
        if (lane == kWaveLanes - 1) {
            S[n] = x;
            bad = bad || x != 1;
        }
        if (bad) flag[0] = 1;
    });
    if (flag[0]) return -1;
    // 2. sizes; leaves know their need
    w.each([&](int lane) {
        for (int i = lane; i < n; i += kWaveLanes) {
            const int target = S[i] + 1;
            int j = i + 1;
            while (j < n) {
                const int v = S[j];
                if (v >= target) break;
                j += target - v;
            }
            size[i] = (j < n ? j : n) - i;
            need[i] = is_leaf(tok[i]) ? 0 : -1;
        }
    });
    // 3. stack need, contiguous blocks from the right (a child follows its
    // parent, so a block's chains settle in one round)
    for (int round = 0; round <= n; ++round) {
        const int f = 1 + (round & 1);
        w.each([&](int lane) {
            if (lane == 0) flag[3 - f] = 0;
            bool pending = false;
            for (int i = wave_block_hi(n, lane) - 1; i >= wave_block_lo(n, lane); --i) {
                if (need[i] >= 0) continue;
                const int t = tok[i];
                const int a = i + 1;
                int r = -1;
                if (!is_binary(t)) {
                    r = need[a];
                } else {
                    const int b = a + size[a];
                    const bool la = is_leaf(tok[a]), lb = is_leaf(tok[b]);
                    const int na = need[a], nb = need[b];
                    if (la && lb) r = 0;
                    else if (la) r = nb;
                    else if (lb) r = na;
                    else if (na >= 0 && nb >= 0) r = na != nb ? (na > nb ? na : nb) : na + 1;
                }
                if (r >= 0) need[i] = r;
                else pending = true;
            }
            if (pending) flag[f] = 1;
        });
        if (!flag[f]) break;
    }
    *stack_need = need[0];
    if (is_leaf(tok[0])) {  // n == 1
        int src;
        int32_t arg;
        leaf_src(tok[0], val[0], nv, src, arg);
        const bool ref = src == SRC_REF && pf;
        out[0] = {0 | (src << 8) | (CL_LOAD << BC_CLS_SHIFT) | (ref ? 1 << BC_PF_SHIFT : 0), arg};
        if (pf)
            for (int j = 0; j < kRefPrefetch; ++j) pf[j] = j == 0 && ref ? arg : -1;
        return 1;
    }
    // 4. instruction counts (prefix W); pointer-jumping state of every node
    auto weight = [&](int i) -> int {
        const int t = tok[i];
        if (is_leaf(t)) return 0;
        if (!is_binary(t)) return 1;
        const int a = i + 1, b = a + size[a];
        return is_leaf(tok[a]) && is_leaf(tok[b]) ? 2 : 1;
    };
    w.each([&](int lane) {
        int x = 0;
        for (int k = wave_block_lo(n, lane); k < wave_block_hi(n, lane); ++k) {
            x += weight(k);
            s.up0[k] = -1;
            s.acc0[k] = 0;
        }
        part[lane] = x;
        for (int k = lane; k < (n + 31) / 32; k += kWaveLanes) s.bits[k] = 0u;
    });
    w.each([&](int lane) {
        int x = 0;
        for (int l = 0; l < lane; ++l) x += part[l];
        for (int k = wave_block_lo(n, lane); k < wave_block_hi(n, lane); ++k) {
            W[k] = x;
            x += weight(k);
        }
        if (lane == kWaveLanes - 1) W[n] = x;
    });
    auto count = [&](int i) { return W[i + size[i]] - W[i]; };
    // first / second evaluated child of node i (-1 when none)
    auto children = [&](int i, int &first, int &second) {
        first = second = -1;
        const int t = tok[i], a = i + 1;
        if (!is_binary(t)) {
            if (!is_leaf(tok[a])) first = a;
            return;
        }
        const int b = a + size[a];
        const bool la = is_leaf(tok[a]), lb = is_leaf(tok[b]);
        if (la && lb) return;
        if (la) first = b;
        else if (lb) first = a;
        else {
            const bool a_first = need[a] > need[b];
            first = a_first ? a : b;
            second = a_first ? b : a;
        }
    };
    w.each([&](int lane) {
        for (int i = lane; i < n; i += kWaveLanes) {
            const int t = tok[i];
            if (is_leaf(t)) continue;
            if (t == OP_STORE && is_leaf(tok[i + 1])) flag[0] = 1;
            int f, g;
            children(i, f, g);
            if (f >= 0) s.up0[f] = i;  // acc0[f] stays 0
            if (g >= 0) {
                s.up0[g] = i;
                s.acc0[g] = count(f);
            }
        }
    });
    if (flag[0]) return -1;
    // 5. subtree starts by pointer jumping, until every node has reached the
    // root (log2 of the height rounds)
    int16_t *acc = s.acc0, *up = s.up0, *acc2 = s.acc1, *up2 = s.up1;
    for (int round = 0, span = 1; span < n; span <<= 1, ++round) {
        const int f = 1 + (round & 1);
        w.each([&](int lane) {
            if (lane == 0) flag[3 - f] = 0;
            bool more = false;
            for (int i = lane; i < n; i += kWaveLanes) {
                const int u = up[i];
                const int u2 = u >= 0 ? up[u] : -1;
                acc2[i] = u >= 0 ? acc[i] + acc[u] : acc[i];
                up2[i] = u2;
                more = more || u2 >= 0;
            }
            if (more) flag[f] = 1;
        });
        int16_t *t0 = acc, *t1 = up;
        acc = acc2;
        up = up2;
        acc2 = t0;
        up2 = t1;
        if (!flag[f]) break;
    }
    // 6. the instructions of operator node i (1, or 2 for a binary operator
    // of two leaves: its LOAD at start), as (position, code, arg).  The
    // instruction a subtree starts with pushes the accumulator unless the
    // subtree starts the program: start > 0 means it is evaluated after a
    // sibling whose value waits on the stack.
    auto node_ins = [&](int i, int *pos, int32_t *code, int32_t *arg) -> int {
        const int t = tok[i];
        const int start = acc[i], at = start + count(i) - 1;
        const int push = start > 0 ? BC_PUSH : 0;
        const int a = i + 1;
        int src;
        int32_t g;
        pos[0] = at;
        if (!is_binary(t)) {
            if (is_leaf(tok[a])) {
                leaf_src(tok[a], val[a], nv, src, g);
                code[0] = t | (src << 8) | push | (CL_UN << BC_CLS_SHIFT);
                arg[0] = g;
            } else if (t == OP_STORE) {
                code[0] = STORE_CODE;
                arg[0] = (int32_t)val[i];
            } else {
                code[0] = t | (SRC_NONE << 8) | (CL_UN << BC_CLS_SHIFT);
                arg[0] = 0;
            }
            return 1;
        }
        const int b = a + size[a];
        const bool la = is_leaf(tok[a]), lb = is_leaf(tok[b]);
        const int bin = CL_BIN << BC_CLS_SHIFT;
        if (la && lb) {
            leaf_src(tok[a], val[a], nv, src, g);
            pos[0] = start;
            code[0] = 0 | (src << 8) | push | (CL_LOAD << BC_CLS_SHIFT);
            arg[0] = g;
            leaf_src(tok[b], val[b], nv, src, g);
            pos[1] = at;
            code[1] = t | (src << 8) | bin;
            arg[1] = g;
            return 2;
        }
        if (la) {
            leaf_src(tok[a], val[a], nv, src, g);
            code[0] = t | (src << 8) | (is_commutative(t) ? 0 : BC_SWAP) | bin;
            arg[0] = g;
        } else if (lb) {
            leaf_src(tok[b], val[b], nv, src, g);
            code[0] = t | (src << 8) | bin;
            arg[0] = g;
        } else {
            const bool a_first = need[a] > need[b];
            code[0] = t | (SRC_POP << 8) | ((a_first && !is_commutative(t)) ? BC_SWAP : 0) | bin;
            arg[0] = 0;
        }
        return 1;
    };
    if (pf) {  // positions of the OP_REF operands
        w.each([&](int lane) {
            for (int i = lane; i < n; i += kWaveLanes) {
                if (is_leaf(tok[i])) continue;
                int pos[2];
                int32_t code[2], arg[2];
                const int k = node_ins(i, pos, code, arg);
                for (int j = 0; j < k; ++j)
                    if (bc_src(code[j]) == SRC_REF) wave_or(&s.bits[pos[j] >> 5], 1u << (pos[j] & 31));
            }
        });
    }
    w.each([&](int lane) {
        for (int i = lane; i < n; i += kWaveLanes) {
            if (is_leaf(tok[i])) continue;
            int pos[2];
            int32_t code[2], arg[2];
            const int k = node_ins(i, pos, code, arg);
            for (int j = 0; j < k; ++j) {
                if (pf && bc_src(code[j]) == SRC_REF) {
                    const int r = bits_below(s.bits, pos[j], kRefPrefetch);
                    if (r < kRefPrefetch) {
                        code[j] |= (r + 1) << BC_PF_SHIFT;
                        pf[r] = arg[j];
                    }
                }
                out[pos[j]] = {code[j], arg[j]};
            }
        }
        if (pf && lane < kRefPrefetch && bits_below(s.bits, W[n], kRefPrefetch) <= lane) pf[lane] = -1;
    });
    return W[n];
}

}  // namespace sr
