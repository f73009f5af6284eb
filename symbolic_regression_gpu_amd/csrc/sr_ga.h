// sr_ga.h -- genetic-algorithm operators on the prefix-token format, one code
// path for host and device, driven by a counter-based RNG (Philox4x32-10).
//
// Semantics restate the reference's host GA (utils.cpp:986-1272):
//   generate  = grow generator (prob_leaf, prob_const_leaf, prob_unary, consts
//               U[const_min, const_max)) bounded by max_depth / a token budget,
//               with the reference's fallback leaf (VAR 0) when the budget runs
//               out under a pending child, so a tree may exceed the budget by
//               up to max_depth tokens;
//   mutate    = replace the subtree at a uniformly drawn node by a fresh tree
//               within max_tokens - (len - old subtree);
//   crossover = splice a uniformly drawn subtree of the right parent into a
//               uniformly drawn node of the left one, <= 32 tries to fit
//               max_tokens, else copy the left parent;
// and the selection of cl_sr_system.cpp:285-330: sort by fitness (NaN last),
// copy the elite fraction, tournament winners by strict '<' on fitness.
//
// Random draws: individual c of generation g draws a private stream
// Philox(key = seed, counter = {block, c, g, domain}); the draw order equals
// the recursion order of the reference generator, so CPU and GPU runs of the
// same (seed, g, fitness vector) produce identical offspring.
#pragma once

#include <stdint.h>

#include "sr_format.h"

namespace sr {

constexpr int GA_MAX_OPS = 32;
constexpr uint32_t GA_DOMAIN_NEXT = 1u;   // offspring of a generation
constexpr uint32_t GA_DOMAIN_INIT = 2u;   // initial population
constexpr uint32_t GA_GEN_INIT = 0xFFFFFFFFu;

struct GaParams {
    int32_t num_vars;
    int32_t max_tokens;
    int32_t max_depth;
    float const_min, const_max;
    float prob_leaf, prob_const_leaf, prob_unary;
    int32_t n_unary, n_binary;
    int32_t unary_ops[GA_MAX_OPS];
    int32_t binary_ops[GA_MAX_OPS];
};

struct SelectParams {
    int32_t tournament;
    float p_mutation;
    int32_t elites;  // number of elite copies (>= 1)
};

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
SR_HD void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                         uint32_t k1, uint32_t out[4])
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

struct Rng {
    uint32_t k0, k1, c1, c2, c3;
    uint32_t word;
    uint32_t buf[4];

    SR_HD void init(uint64_t seed, uint32_t stream, uint32_t gen, uint32_t domain)
    {
        k0 = (uint32_t)seed;
        k1 = (uint32_t)(seed >> 32);
        c1 = stream;
        c2 = gen;
        c3 = domain;
        word = 0;
    }
    // word w of the stream is output w & 3 of Philox block w >> 2
    SR_HD uint32_t next()
    {
        const uint32_t j = word & 3u;
        if (j == 0u) philox4x32_10(word >> 2, c1, c2, c3, k0, k1, buf);
        ++word;
        // selects, not a dynamic index: keeps buf in registers on the device
        return j == 0u ? buf[0] : (j == 1u ? buf[1] : (j == 2u ? buf[2] : buf[3]));
    }
    // [lo, hi): lo + (hi - lo) * u with u = top 24 bits / 2^24; each operation
    // rounds once (the build disables contraction).
    SR_HD float uniform(float lo, float hi)
    {
        const float u = (float)(next() >> 8) * (1.0f / 16777216.0f);
        const float span = hi - lo;
        const float t = span * u;
        return lo + t;
    }
    // [lo, hi] by multiply-shift; no draw when hi <= lo (utils.cpp:973-979).
    SR_HD int randint(int lo, int hi)
    {
        if (hi <= lo) return lo;
        const uint32_t range = (uint32_t)(hi - lo) + 1u;
        return lo + (int)(((uint64_t)next() * range) >> 32);
    }
    // uniform(0, 1) < p with t = chance_threshold(p), in integers (the same
    // draw, the same outcome)
    SR_HD bool chance(uint32_t t) { return (next() >> 8) < t; }
};

#if defined(__HIPCC__)
// The same stream for a wave whose 64 lanes all consume it alike: lane k
// holds Philox block base / 4 + k, so one parallel round of Philox gives the
// next 256 words, and a word is a read of one lane.  Must be used with every
// lane of the wave active, in uniform control flow.
struct WaveRng {
    uint32_t k0, k1, c1, c2, c3;
    uint32_t word, base;
    uint32_t b0, b1, b2, b3;

    __device__ void init(uint64_t seed, uint32_t stream, uint32_t gen, uint32_t domain)
    {
        k0 = (uint32_t)seed;
        k1 = (uint32_t)(seed >> 32);
        c1 = stream;
        c2 = gen;
        c3 = domain;
        word = 0;
        fill();
    }
    __device__ void fill()
    {
        base = word & ~3u;
        uint32_t o[4];
        philox4x32_10((base >> 2) + (uint32_t)(threadIdx.x & 63), c1, c2, c3, k0, k1, o);
        b0 = o[0];
        b1 = o[1];
        b2 = o[2];
        b3 = o[3];
    }
    __device__ uint32_t next()
    {
        if (word - base >= 256u) fill();
        const uint32_t idx = word - base;
        ++word;
        const int src = (int)(idx >> 2);
        switch (idx & 3u) {
        case 0: return (uint32_t)__builtin_amdgcn_readlane((int)b0, src);
        case 1: return (uint32_t)__builtin_amdgcn_readlane((int)b1, src);
        case 2: return (uint32_t)__builtin_amdgcn_readlane((int)b2, src);
        default: return (uint32_t)__builtin_amdgcn_readlane((int)b3, src);
        }
    }
    __device__ float uniform(float lo, float hi)
    {
        const float u = (float)(next() >> 8) * (1.0f / 16777216.0f);
        const float span = hi - lo;
        const float t = span * u;
        return lo + t;
    }
    __device__ int randint(int lo, int hi)
    {
        if (hi <= lo) return lo;
        const uint32_t range = (uint32_t)(hi - lo) + 1u;
        return lo + (int)(((uint64_t)next() * range) >> 32);
    }
    __device__ bool chance(uint32_t t) { return (next() >> 8) < t; }
    // consume a word whose value is computed elsewhere (word_at)
    __device__ void skip() { ++word; }
    // word w of the stream, computed on its own (any lane, any w)
    __device__ uint32_t word_at(uint32_t w) const
    {
        uint32_t o[4];
        philox4x32_10(w >> 2, c1, c2, c3, k0, k1, o);
        const uint32_t j = w & 3u;
        return j == 0u ? o[0] : (j == 1u ? o[1] : (j == 2u ? o[2] : o[3]));
    }
};
#endif

// uniform(0, 1) is exactly m / 2^24 (m = next() >> 8; the product by 1 and
// the sum with 0 are exact), so uniform(0, 1) < p  <=>  m < p * 2^24  <=>
// m < ceil(p * 2^24) for p in (0, 1); never for p <= 0 or NaN, always for
// p >= 1.
SR_HD uint32_t chance_threshold(float p)
{
    if (!(p > 0.0f)) return 0u;
    if (p >= 1.0f) return 1u << 24;
    const float x = p * 16777216.0f;  // exact: a power of two
    const uint32_t f = (uint32_t)x;
    return (float)f == x ? f : f + 1u;
}

// opcodes.h:42-78 arity: unary 10..27 (and the clamped exp 29), everything
// else binary; leaves have none.
SR_HD int ga_arity(int t)
{
    if (t == TOK_CONST || t == TOK_VAR) return 0;
    if ((t >= 10 && t <= 27) || t == OP_EXP_CLAMPED) return 1;
    return 2;
}

// utils.cpp:1110-1129 subtree extent, iteratively.
SR_HD int ga_subtree_size(const int32_t *tok, int len, int start)
{
    if (start < 0 || start >= len) return 0;
    int need = 1, pos = start;
    while (need > 0 && pos < len) {
        need += ga_arity(tok[pos]) - 1;
        ++pos;
    }
    return pos - start;
}

// Pending-child stack of the grow generator (entries: depth, or -1 for the
// top-level call).
constexpr int GA_GEN_STACK = 2 * 64 + 4;

struct GaArrayStack {
    int16_t a[GA_GEN_STACK];
    int sp = 0;
    SR_HD void push(int v) { a[sp++] = (int16_t)v; }
    SR_HD int pop() { return a[--sp]; }
};
struct GaParamOps {
    const GaParams &p;
    SR_HD int unary(int i) const { return p.unary_ops[i]; }
    SR_HD int binary(int i) const { return p.binary_ops[i]; }
};

// Grow generator (utils.cpp:995-1108) writing at most cap tokens into
// tok/val.  Returns the length, or -1 if cap would be exceeded.  With store
// false it only draws (the lanes of a wave that run the same generator and
// leave the writes to one lane).  Stack and Ops hold the pending children and
// the operator tables (GaArrayStack / GaParamOps, or their per-wave forms).
template <class R, class Stack, class Ops>
SR_HD __attribute__((always_inline)) int ga_generate_with(const GaParams &p, int budget, R &r, int32_t *tok, float *val, int cap, bool store,
                           Stack &stack, const Ops &ops)
{
    if (budget <= 0) return 0;
    int left = p.max_tokens < budget ? p.max_tokens : budget;
    const int nvars = p.num_vars > 1 ? p.num_vars : 1;
    const uint32_t t_leaf = chance_threshold(p.prob_leaf), t_const = chance_threshold(p.prob_const_leaf),
                   t_unary = chance_threshold(p.prob_unary);
    int n = 0;
    stack.push(-1);
    while (stack.sp > 0) {
        int d = stack.pop();
        const bool top = d < 0;
        if (top) d = 0;
        if (left <= 0) {
            if (!top) {  // budget exhausted under a pending child: fallback leaf
                if (n >= cap) return -1;
                if (store) {
                    tok[n] = TOK_VAR;
                    val[n] = 0.0f;
                }
                ++n;
            }
            continue;
        }
        bool leaf = d >= p.max_depth || left == 1;
        if (!leaf && r.chance(t_leaf)) leaf = true;
        if (!leaf && !(left >= 2)) leaf = true;
        if (leaf) {
            if (n >= cap) return -1;
            bool is_const = r.chance(t_const);
            const float lv = is_const ? r.uniform(p.const_min, p.const_max) : (float)r.randint(0, nvars - 1);
            if (store) {
                tok[n] = is_const ? TOK_CONST : TOK_VAR;
                val[n] = lv;
            }
            ++n;
            --left;
            continue;
        }
        const bool unary = r.chance(left >= 3 ? t_unary : 1u << 24);
        int op, arity;
        if (unary) {
            op = ops.unary(r.randint(0, p.n_unary - 1));
            arity = 1;
        } else {
            op = ops.binary(r.randint(0, p.n_binary - 1));
            arity = 2;
        }
        if (n >= cap) return -1;
        if (store) {
            tok[n] = op;
            val[n] = 0.0f;
        }
        ++n;
        --left;
        if (stack.sp + arity > GA_GEN_STACK) return -1;
        for (int k = 0; k < arity; ++k) stack.push(d + 1);
    }
    if (n == 0) {
        if (store) {
            tok[0] = TOK_CONST;
            val[0] = 0.0f;
        }
        n = 1;
    }
    return n;
}

SR_HD int ga_generate(const GaParams &p, int budget, Rng &r, int32_t *tok, float *val, int cap, bool store = true)
{
    GaArrayStack stack;
    return ga_generate_with(p, budget, r, tok, val, cap, store, stack, GaParamOps{p});
}

SR_HD int ga_copy(const int32_t *st, const float *sv, int len, int32_t *dt, float *dv)
{
    for (int k = 0; k < len; ++k) {
        dt[k] = st[k];
        dv[k] = sv[k];
    }
    return len;
}

// mutate_expression (utils.cpp:1163-1204); child written in place of a
// splice: prefix, fresh subtree generated straight into the child, suffix.
SR_HD int ga_mutate(const GaParams &p, const int32_t *pt, const float *pv, int plen, Rng &r,
                    int32_t *ct, float *cv, int cap)
{
    if (plen <= 0) return ga_generate(p, p.max_tokens, r, ct, cv, cap);
    const int idx = r.randint(0, plen - 1);
    int old = ga_subtree_size(pt, plen, idx);
    if (old <= 0) old = 1;
    const int base = plen - old;
    const int max_new = p.max_tokens - base;
    if (max_new <= 0) return ga_copy(pt, pv, plen, ct, cv);
    ga_copy(pt, pv, idx, ct, cv);
    const int tail = plen - (idx + old);
    const int room = cap - idx - tail;
    const int nl = ga_generate(p, max_new, r, ct + idx, cv + idx, room);
    if (nl < 0) return -1;
    ga_copy(pt + idx + old, pv + idx + old, tail, ct + idx + nl, cv + idx + nl);
    return idx + nl + tail;
}

// crossover_expressions (utils.cpp:1206-1272).
SR_HD int ga_crossover(const GaParams &p, const int32_t *lt, const float *lv, int llen,
                       const int32_t *rt, const float *rv, int rlen, Rng &r, int32_t *ct,
                       float *cv, int cap)
{
    if (llen <= 0 || rlen <= 0) return ga_copy(lt, lv, llen, ct, cv);
    const int li = r.randint(0, llen - 1);
    int lsub = ga_subtree_size(lt, llen, li);
    if (lsub <= 0) lsub = 1;
    const int attempts = rlen < 32 ? rlen : 32;
    for (int a = 0; a < attempts; ++a) {
        const int ri = r.randint(0, rlen - 1);
        const int rsub = ga_subtree_size(rt, rlen, ri);
        if (rsub <= 0) continue;
        const int total = llen - lsub + rsub;
        if (total > p.max_tokens) continue;
        if (total > cap) return -1;
        ga_copy(lt, lv, li, ct, cv);
        ga_copy(rt + ri, rv + ri, rsub, ct + li, cv + li);
        ga_copy(lt + li + lsub, lv + li + lsub, llen - li - lsub, ct + li + rsub, cv + li + rsub);
        return total;
    }
    return ga_copy(lt, lv, llen, ct, cv);
}

// Strict weak order of cl_sr_system.cpp:287-291 made total: ascending
// fitness, NaN last, ties by index.
SR_HD bool fitness_before(float fa, int a, float fb, int b)
{
    const bool na = fa != fa, nb = fb != fb;
    if (na != nb) return nb;
    if (!na && fa != fb) return fa < fb;
    return a < b;
}

// Tournament over the sorted population (cl_sr_system.cpp:305-313).
SR_HD int ga_tournament(Rng &r, const float *fit, const int32_t *order, int pop, int t)
{
    int best = r.randint(0, pop - 1);
    for (int k = 0; k < t - 1; ++k) {
        const int cand = r.randint(0, pop - 1);
        if (fit[order[cand]] < fit[order[best]]) best = cand;
    }
    return best;
}

// One offspring of generation gen: elite copy or tournament + variation.
// Population layout: individual i at tok/val + i*cap, length len[i].
SR_HD int ga_child(const GaParams &p, const SelectParams &s, int pop, int cap, const int32_t *tok,
                   const float *val, const int32_t *len, const float *fit, const int32_t *order,
                   uint64_t seed, uint32_t gen, int c, int32_t *ct, float *cv)
{
    if (c < s.elites) {
        const int e = order[c];
        return ga_copy(tok + (size_t)e * cap, val + (size_t)e * cap, len[e], ct, cv);
    }
    Rng r;
    r.init(seed, (uint32_t)c, gen, GA_DOMAIN_NEXT);
    const int b1 = order[ga_tournament(r, fit, order, pop, s.tournament)];
    if (r.chance(chance_threshold(s.p_mutation)))
        return ga_mutate(p, tok + (size_t)b1 * cap, val + (size_t)b1 * cap, len[b1], r, ct, cv, cap);
    const int b2 = order[ga_tournament(r, fit, order, pop, s.tournament)];
    return ga_crossover(p, tok + (size_t)b1 * cap, val + (size_t)b1 * cap, len[b1],
                        tok + (size_t)b2 * cap, val + (size_t)b2 * cap, len[b2], r, ct, cv, cap);
}

// ---- one wave per child ----------------------------------------------------
// ga_child's decisions (the same draws in the same order, every lane running
// them alike), with the parents' tokens staged in LDS for the subtree scans
// and the copies spread over the lanes; generated tokens are written by lane
// 0.  Offspring are identical to ga_child's.
__device__ inline void ga_wave_copy(const int32_t *st, const float *sv, int len, int32_t *dt, float *dv, int lane)
{
    for (int k = lane; k < len; k += 64) {
        dt[k] = st[k];
        dv[k] = sv[k];
    }
}

// A parent's subtree extents: the wave computes S[k] = sum over tokens before
// k of (1 - ga_arity) into LDS; the subtree at `start` ends at the first
// j > start with S[j] = S[start] + 1 (the end ga_subtree_size's count
// reaches), and since S rises by at most 1 per token the search may skip
// target - S[j] - 1 positions.
__device__ inline void ga_wave_stage(const int32_t *src, int len, int32_t *S, int lane)
{
    len = len > 0 ? len : 0;
    const int per = (len + 63) / 64;
    const int lo = min(len, lane * per), hi = min(len, lo + per);
    int d = 0;
    for (int k = lo; k < hi; ++k) d += 1 - ga_arity(src[k]);
    int inc = d;
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(inc, o, 64);
        if (lane >= o) inc += u;
    }
    int x = inc - d;
    for (int k = lo; k < hi; ++k) {
        S[k] = x;
        x += 1 - ga_arity(src[k]);
    }
    if (lane == 63) S[len] = x;
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): this wave's LDS writes done
    __builtin_amdgcn_wave_barrier();
}

__device__ inline int ga_subtree_size_s(const int32_t *S, int len, int start)
{
    if (start < 0 || start >= len) return 0;
    const int target = S[start] + 1;
    int j = start + 1;
    while (j < len) {
        const int v = S[j];
        if (v >= target) break;
        j += target - v;
    }
    return min(j, len) - start;
}

// ga_tournament with the candidates drawn first (the same draws in the same
// order) and their fitness loaded 8 at a time, so a tournament waits for a
// few memory round trips instead of two per candidate.
template <class R>
__device__ inline int ga_tournament_wave(R &r, const float *fit, const int32_t *order, int pop, int t)
{
    constexpr int B = 8;
    int best = r.randint(0, pop - 1);
    float fb = fit[order[best]];
    for (int k0 = 0; k0 < t - 1; k0 += B) {
        const int m = min(B, t - 1 - k0);
        int cand[B];
#pragma unroll
        for (int k = 0; k < B; ++k) cand[k] = k < m ? r.randint(0, pop - 1) : cand[0];
        int who[B];
#pragma unroll
        for (int k = 0; k < B; ++k) who[k] = order[cand[k]];
        float f[B];
#pragma unroll
        for (int k = 0; k < B; ++k) f[k] = fit[who[k]];
#pragma unroll
        for (int k = 0; k < B; ++k)
            if (k < m && f[k] < fb) {
                best = cand[k];
                fb = f[k];
            }
    }
    return best;
}

// LDS ints per wave of k_ga_next: two prefix rows and the generator stack.
SR_HD int ga_wave_lds_ints(int cap) { return 2 * (cap + 1) + (GA_GEN_STACK + 1) / 2; }

// The generator's stack in LDS shared by the wave's lanes (every lane pushes
// the same values).
struct GaLdsStack {
    int16_t *a;
    int sp = 0;
    __device__ void push(int v) { a[sp++] = (int16_t)v; }
    // the value is the same in every lane: saying so keeps the generator's
    // decisions on the scalar unit
    __device__ int pop() { return __builtin_amdgcn_readfirstlane(a[--sp]); }
};

// ga_generate for a wave, in two passes.  The draws that decide the shape
// (leaf? constant? unary?) are made in order by the wave; a token's value
// draw (constant, variable index, operator index) is only counted, its word
// index kept in plan[k] with the token kind in bits 28..31.  The lanes then
// compute every token's value from its word at once.  Same draws, same
// words, same tokens as ga_generate (utils.cpp:995-1108).
enum { GEN_FALLBACK = 0, GEN_CONST, GEN_VAR, GEN_UNARY, GEN_BINARY };

__device__ inline int ga_generate_wave(const GaParams &p, int budget, WaveRng &r, int32_t *tok, float *val,
                                       int cap, GaLdsStack &stack, int32_t *plan, int lane)
{
    if (budget <= 0) return 0;
    int left = p.max_tokens < budget ? p.max_tokens : budget;
    const int nvars = p.num_vars > 1 ? p.num_vars : 1;
    const uint32_t t_leaf = chance_threshold(p.prob_leaf), t_const = chance_threshold(p.prob_const_leaf),
                   t_unary = chance_threshold(p.prob_unary);
    int n = 0;
    stack.push(-1);
    while (stack.sp > 0) {
        int d = stack.pop();
        const bool top = d < 0;
        if (top) d = 0;
        if (left <= 0) {
            if (!top) {  // budget exhausted under a pending child: fallback leaf
                if (n >= cap) return -1;
                plan[n++] = GEN_FALLBACK << 28;
            }
            continue;
        }
        bool leaf = d >= p.max_depth || left == 1;
        if (!leaf && r.chance(t_leaf)) leaf = true;
        if (leaf) {
            if (n >= cap) return -1;
            const bool is_const = r.chance(t_const);
            plan[n++] = ((is_const ? GEN_CONST : GEN_VAR) << 28) | (int32_t)r.word;
            if (is_const || nvars > 1) r.skip();
            --left;
            continue;
        }
        const bool unary = r.chance(left >= 3 ? t_unary : 1u << 24);
        const int32_t w = (int32_t)r.word;
        if (unary ? p.n_unary > 1 : p.n_binary > 1) r.skip();
        if (n >= cap) return -1;
        plan[n++] = ((unary ? GEN_UNARY : GEN_BINARY) << 28) | w;
        --left;
        if (stack.sp + (unary ? 1 : 2) > GA_GEN_STACK) return -1;
        stack.push(d + 1);
        if (!unary) stack.push(d + 1);
    }
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): the plan is in LDS
    __builtin_amdgcn_wave_barrier();
    for (int k = lane; k < n; k += 64) {
        const int32_t e = plan[k];
        const int kind = (e >> 28) & 15;
        const uint32_t m = kind == GEN_FALLBACK ? 0u : r.word_at((uint32_t)(e & 0x0fffffff));
        int t = TOK_VAR;
        float v = 0.0f;
        if (kind == GEN_CONST) {  // uniform(const_min, const_max)
            t = TOK_CONST;
            const float u = (float)(m >> 8) * (1.0f / 16777216.0f);
            const float span = p.const_max - p.const_min;
            const float tt = span * u;
            v = p.const_min + tt;
        } else if (kind == GEN_VAR) {  // randint(0, nvars - 1)
            v = nvars > 1 ? (float)(int)(((uint64_t)m * (uint32_t)nvars) >> 32) : 0.0f;
        } else if (kind == GEN_UNARY) {
            t = p.unary_ops[p.n_unary > 1 ? (int)(((uint64_t)m * (uint32_t)p.n_unary) >> 32) : 0];
        } else if (kind == GEN_BINARY) {
            t = p.binary_ops[p.n_binary > 1 ? (int)(((uint64_t)m * (uint32_t)p.n_binary) >> 32) : 0];
        }
        tok[k] = t;
        val[k] = v;
    }
    if (n == 0) {
        if (lane == 0) {
            tok[0] = TOK_CONST;
            val[0] = 0.0f;
        }
        n = 1;
    }
    return n;
}

__device__ inline int ga_child_wave(const GaParams &p, const SelectParams &s, int pop, int cap,
                                    const int32_t *tok, const float *val, const int32_t *len, const float *fit,
                                    const int32_t *order, uint64_t seed, uint32_t gen, int c, int32_t *ct,
                                    float *cv, int32_t *l1, int32_t *l2, int16_t *gstack, int lane)
{
    if (c < s.elites) {
        const int e = order[c];
        ga_wave_copy(tok + (size_t)e * cap, val + (size_t)e * cap, len[e], ct, cv, lane);
        return len[e];
    }
    WaveRng r;
    r.init(seed, (uint32_t)c, gen, GA_DOMAIN_NEXT);
    const int b1 = order[ga_tournament_wave(r, fit, order, pop, s.tournament)];
    const int32_t *pt = tok + (size_t)b1 * cap;
    const float *pv = val + (size_t)b1 * cap;
    const int plen = len[b1];
    ga_wave_stage(pt, plen, l1, lane);
    if (r.chance(chance_threshold(s.p_mutation))) {  // ga_mutate
        if (plen <= 0) {
            GaLdsStack st{gstack};
            return ga_generate_wave(p, p.max_tokens, r, ct, cv, cap, st, l2, lane);
        }
        const int idx = r.randint(0, plen - 1);
        int old = ga_subtree_size_s(l1, plen, idx);
        if (old <= 0) old = 1;
        const int base = plen - old;
        const int max_new = p.max_tokens - base;
        if (max_new <= 0) {
            ga_wave_copy(pt, pv, plen, ct, cv, lane);
            return plen;
        }
        ga_wave_copy(pt, pv, idx, ct, cv, lane);
        const int tail = plen - (idx + old);
        const int room = cap - idx - tail;
        GaLdsStack st{gstack};
        const int nl = ga_generate_wave(p, max_new, r, ct + idx, cv + idx, room, st, l2, lane);
        if (nl < 0) return -1;
        ga_wave_copy(pt + idx + old, pv + idx + old, tail, ct + idx + nl, cv + idx + nl, lane);
        return idx + nl + tail;
    }
    const int b2 = order[ga_tournament_wave(r, fit, order, pop, s.tournament)];
    const int32_t *rt = tok + (size_t)b2 * cap;
    const float *rv = val + (size_t)b2 * cap;
    const int rlen = len[b2];
    ga_wave_stage(rt, rlen, l2, lane);
    // ga_crossover
    if (plen <= 0 || rlen <= 0) {
        ga_wave_copy(pt, pv, plen, ct, cv, lane);
        return plen;
    }
    const int li = r.randint(0, plen - 1);
    int lsub = ga_subtree_size_s(l1, plen, li);
    if (lsub <= 0) lsub = 1;
    const int attempts = rlen < 32 ? rlen : 32;
    for (int a = 0; a < attempts; ++a) {
        const int ri = r.randint(0, rlen - 1);
        const int rsub = ga_subtree_size_s(l2, rlen, ri);
        if (rsub <= 0) continue;
        const int total = plen - lsub + rsub;
        if (total > p.max_tokens) continue;
        if (total > cap) return -1;
        ga_wave_copy(pt, pv, li, ct, cv, lane);
        ga_wave_copy(rt + ri, rv + ri, rsub, ct + li, cv + li, lane);
        ga_wave_copy(pt + li + lsub, pv + li + lsub, plen - li - lsub, ct + li + rsub, cv + li + rsub, lane);
        return total;
    }
    ga_wave_copy(pt, pv, plen, ct, cv, lane);
    return plen;
}

}  // namespace sr
