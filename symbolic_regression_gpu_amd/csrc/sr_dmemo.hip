// sr_dmemo.hip -- the generational subtree cache planned on the device.
//
// No counterpart in the reference: its cache (cl_gpu_eval.cu:294-329,
// 384-397) is planned on the host, keeps the first subtrees it detects and
// never evicts.  Here a generation of the device-resident GA population is
// planned by kernels, so evaluate -> fitness -> select -> vary runs with no
// host work and is captured as one hipGraph per generation.
//
// Per generation (E expressions of stride cap, positions p = e * cap + q):
//   1. sizes and 64-bit keys (span hashes) of every subtree;
//   2. every subtree of >= min_size tokens is looked up among the slots; a
//      hit (key and token body equal) will be read from its column;
//   3. positions outside hits are "executed"; executed subtrees without a
//      column are counted by key;
//   4. slots whose subtree the population no longer holds are evicted;
//   5. a subtree executed at least twice ("shared") gets a slot: a shared
//      subtree with an occurrence inside no other shared occurrence is a
//      root: launch 1 evaluates it once (from its first such occurrence) and
//      writes its column; a shared subtree only ever met inside others is
//      written by an OP_STORE inside the launch-1 program of the root that
//      holds its first occurrence;
//   6. launch 2 runs the expressions with every hit and every shared subtree
//      read as an OP_REF; an expression that is itself a root takes its MSE
//      from launch 1, one whose whole tree has a column and a remembered MSE
//      runs nothing.
// Columns hold exactly what the interpreter computes for a subtree (the same
// bytecode, the same operands), so every MSE equals the uncached batched one
// bit for bit.  Slot numbers follow the first positions of the new subtrees,
// so plans are deterministic.
#include <algorithm>
#include <climits>

#include "sr_compile_wave.h"
#include "sr_ctx.h"
#include "sr_ga_launch.h"
#include "sr_kernels.h"
#include "sr_launch.h"

namespace sr {

namespace {

constexpr int kThreads = 256;
constexpr int32_t kNone = INT_MAX;
constexpr int32_t kFromLaunch1 = 0x40000000;  // e_src: launch-1 program r is kFromLaunch1 | r
constexpr int32_t kHitPending = -6;            // hit: kHitPending - slot awaits its body compare
constexpr int kQueueCounters = 64;             // queue counters per launch (c.d_counters holds two sets)

// counters (int32): generation, free slots, launch-1 programs, nested slots
// given, launch-2 programs, roots found, nested found, status
enum { C_GEN = 0, C_FREE, C_L1, C_NEST, C_L2, C_ROOTS, C_NESTED, C_STATUS, C_COUNT };
// statistics (int64), accumulated over generations
enum { S_LOGICAL = 0, S_EXEC, S_REFS, S_STORES, S_OPS, S_L1, S_L2, S_LIVE, S_GENS, S_COUNT };
// Counters every program or expression adds to (S_LOGICAL .. S_OPS) are
// spread over kStatSpread copies after the S_COUNT totals (stats[S_COUNT +
// copy * S_COUNT + k]), so thousands of blocks do not queue on one address;
// the host sums the copies.
constexpr int kStatSpread = 64;
SR_DEV void stat_add(long long *stats, int who, int k, long long v)
{
    atomicAdd((unsigned long long *)&stats[S_COUNT + (who & (kStatSpread - 1)) * S_COUNT + k], (unsigned long long)v);
}

SR_DEV uint64_t mix64(uint64_t x)
{
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33;
    return x;
}

SR_DEV int tok_arity(int t) { return t <= 0 ? 0 : (t < 10 ? 2 : 1); }

struct DmArgs {
    const int32_t *tok;
    const float *val;
    const int32_t *len;
    int E, cap, PC, L1, C, H, H2, ms;
    int32_t *sz;
    uint64_t *hh;
    int32_t *hit, *sh, *anc, *pslot, *cidx, *flag, *rank;
    uint64_t *s_key;
    int32_t *s_len, *s_tok, *s_src;
    float *s_val;
    uint32_t *s_seen;
    int32_t *s_live;
    float *s_mse;
    long long *s_cnt;
    int32_t *s_has, *s_free;
    uint64_t *ht_key;
    int32_t *ht_slot;
    uint64_t *ct_key;
    int32_t *ct_cnt, *ct_rep, *ct_root, *ct_slot, *ct_wr, *ct_off, *ct_prog;
    int32_t *p1_slot, *p1_pos, *p1_need;
    int32_t *e_src;
    Ins *p1_ins, *p2_ins;
    int32_t *p1_plen, *p2_plen, *p1_pf, *p2_pf;
    float *mse1, *fit;
    long long *cnt1;
    int32_t *ctr;
    long long *stats;
};

// Token bodies compared 8 tokens at a time: the loads of a chunk are
// independent (indexes clamped to the body), so a long body costs one memory
// latency per chunk instead of one per token.
constexpr int kBodyChunk = 8;

SR_DEV bool same_tokens(const int32_t *__restrict__ t1, const float *__restrict__ v1,
                        const int32_t *__restrict__ t2, const float *__restrict__ v2, int n)
{
    bool same = true;
    for (int j0 = 0; j0 < n && same; j0 += kBodyChunk) {
        const int last = n - 1 - j0;
        int32_t x[kBodyChunk], y[kBodyChunk];
        uint32_t u[kBodyChunk], w[kBodyChunk];
#pragma unroll
        for (int k = 0; k < kBodyChunk; ++k) {
            const int j = j0 + min(k, last);
            x[k] = t1[j];
            y[k] = t2[j];
            u[k] = __float_as_uint(v1[j]);
            w[k] = __float_as_uint(v2[j]);
        }
#pragma unroll
        for (int k = 0; k < kBodyChunk; ++k) same = same && x[k] == y[k] && u[k] == w[k];
    }
    return same;
}

SR_DEV bool same_body(const DmArgs &a, int64_t p, int64_t r, int n)
{
    return same_tokens(a.tok + p, a.val + p, a.tok + r, a.val + r, n);
}

// Exclusive prefix of a per-lane value over the wave's lanes, through LDS
// (part: kWaveLanes ints; every lane must call it).
SR_DEV int lane_prefix_sum(int32_t *part, int v, int lane)
{
    part[lane] = v;
    __syncthreads();
    int x = 0;
    for (int k = 0; k < lane; ++k) x += part[k];
    __syncthreads();
    return x;
}

// Wave-cooperative body compares: candidate k compares n[k] tokens of the
// population at position p[k] with a second body (body2(r[k]) gives its
// tokens and values); mism[k] = 1 where they differ.  Every body is cut into
// chunks of kCoopChunk tokens and the chunks of all candidates are spread
// over the lanes, so a long body costs no more round trips than a short one.
// LDS: p, r, n, off, mism [nc]; part [kWaveLanes].  Every lane must call it.
constexpr int kCoopChunk = 16;

template <class Body2>
SR_DEV void wave_compare(const DmArgs &a, int nc, const int32_t *cp, const int32_t *cr, const int32_t *cn,
                         int32_t *off, int32_t *mism, int32_t *part, int lane, Body2 body2)
{
    const int lo = wave_block_lo(nc, lane), hi = wave_block_hi(nc, lane);
    int x = 0;
    for (int k = lo; k < hi; ++k) x += (cn[k] + kCoopChunk - 1) / kCoopChunk;
    int o = lane_prefix_sum(part, x, lane);
    for (int k = lo; k < hi; ++k) {
        off[k] = o;
        mism[k] = 0;
        o += (cn[k] + kCoopChunk - 1) / kCoopChunk;
    }
    const int total = __shfl(o, kWaveLanes - 1, kWaveLanes);
    __syncthreads();
    for (int c = lane; c < total; c += kWaveLanes) {
        int l = 0, h = nc - 1;  // the last candidate whose chunks start at or before c
        while (l < h) {
            const int mid = (l + h + 1) >> 1;
            if (off[mid] <= c) l = mid;
            else h = mid - 1;
        }
        const int j0 = (c - off[l]) * kCoopChunk;
        const int32_t *t2;
        const float *v2;
        body2(cr[l], t2, v2);
        const int64_t p = cp[l];
        if (!same_tokens(a.tok + p + j0, a.val + p + j0, t2 + j0, v2 + j0, min(kCoopChunk, cn[l] - j0)))
            mism[l] = 1;
    }
    __syncthreads();
}

// ---- per generation: reset (counters, per-position and per-key state) -----------
__global__ void k_dm_clear(DmArgs a)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0) {
        a.ctr[C_GEN] += 1;
        for (int k = 1; k < C_COUNT; ++k) a.ctr[k] = 0;
        a.stats[S_GENS] += 1;
    }
    if (p < (int64_t)a.E * a.cap) {
        a.hit[p] = -3;
        a.sh[p] = 0;
        a.anc[p] = kNone;
        a.pslot[p] = -1;
        a.cidx[p] = -1;
        a.flag[p] = 0;
    }
    if (p < a.L1) {
        a.p1_plen[p] = 0;
        a.p1_need[p] = 0;
    }
    if (p < a.E) a.p2_plen[p] = 0;
    if (p < a.H2) {
        a.ct_key[p] = 0;
        a.ct_cnt[p] = 0;
        a.ct_rep[p] = kNone;
        a.ct_root[p] = kNone;
        a.ct_slot[p] = -1;
        a.ct_wr[p] = -1;
        a.ct_off[p] = 0;
        a.ct_prog[p] = -1;
    }
}

// ---- lookup of a subtree among the slots ------------------------------------------
// hit value: slot (>= 0, body verified), -1 no key, -2 another body's key, -3
// not looked up, kHitPending - slot: key and length match, body compared
// later (by k_dm_count's wave).
SR_DEV int lookup(const DmArgs &a, uint64_t h, int n)
{
    for (uint32_t i = (uint32_t)mix64(h) & (a.H - 1);; i = (i + 1) & (a.H - 1)) {
        const uint64_t k = a.ht_key[i];
        if (k == 0) return -1;
        if (k != h) continue;
        const int s = a.ht_slot[i];
        return a.s_len[s] == n ? kHitPending - s : -2;
    }
}

// ---- 1. sizes, subtree keys and the lookup of every subtree of >= min_size
// tokens (one wave per expression) ------------------------------------------------
// A subtree is the token span [i, i + size) of the prefix stream.  With
// S[k] = sum over tokens before k of (1 - arity), its end is the first
// j > i with S[j] = S[i] + 1; S rises by at most 1 per token, so a search
// from S[j] < target may skip target - S[j] - 1 positions.  Its key is a
// polynomial hash of the span: every token gets a 64-bit code, weighted by
// Binv^k (B odd, Binv its inverse mod 2^64); the prefix sums U of the
// weighted codes give sum_k code_k Binv^(k-i) = (U[end] - U[i]) * B^i,
// which depends only on the span's contents, finalised with its length.
// 0 is reserved for empty table entries.  Both prefixes are wave scans, so
// nothing here is serial over the tokens.
struct HashPowers {
    uint64_t pw[MAX_EXPR_LEN + 1], ipw[MAX_EXPR_LEN + 1];
};
constexpr uint64_t kHashBase = 0xd6e8feb86659fd93ULL;  // odd

constexpr HashPowers make_hash_powers()
{
    HashPowers h{};
    uint64_t inv = kHashBase;  // Newton: each step doubles the correct low bits
    for (int k = 0; k < 6; ++k) inv *= 2 - kHashBase * inv;
    uint64_t p = 1, q = 1;
    for (int k = 0; k <= MAX_EXPR_LEN; ++k) {
        h.pw[k] = p;
        h.ipw[k] = q;
        p *= kHashBase;
        q *= inv;
    }
    return h;
}
static_assert(make_hash_powers().pw[5] * make_hash_powers().ipw[5] == 1, "hash base inverse");
__constant__ HashPowers kHashPowers = make_hash_powers();

SR_DEV uint64_t token_code(int t, uint32_t v)
{
    return mix64((uint64_t)(uint32_t)t * 0x9e3779b97f4a7c15ULL ^ mix64((uint64_t)v + 0x632be59bd9b4e019ULL));
}

template <class T>
SR_DEV T wave_excl_scan(T v, int lane)
{
    T inc = v;
    for (int d = 1; d < 64; d <<= 1) {
        const T u = __shfl_up(inc, d, 64);
        if (lane >= d) inc += u;
    }
    return inc - v;
}

__global__ __launch_bounds__(64) void k_dm_hash(DmArgs a)
{
    extern __shared__ unsigned char lds[];
    uint64_t *lu = (uint64_t *)lds;            // [cap + 1] U prefix
    int32_t *ls = (int32_t *)(lu + a.cap + 1);  // [cap + 1] S prefix
    const int e = blockIdx.x, lane = threadIdx.x;
    const int64_t b = (int64_t)e * a.cap;
    const int L = max(0, min(a.len[e], a.cap));
    // contiguous blocks per lane for the prefixes
    const int per = (L + 63) / 64;
    const int lo = min(L, lane * per), hi = min(L, lo + per);
    int ds = 0;
    uint64_t us = 0;
    for (int k = lo; k < hi; ++k) {
        const int t = a.tok[b + k];
        const uint64_t w = token_code(t, __float_as_uint(a.val[b + k])) * kHashPowers.ipw[k];
        ls[k] = 1 - tok_arity(t);
        lu[k] = w;
        ds += 1 - tok_arity(t);
        us += w;
    }
    int dx = wave_excl_scan(ds, lane);
    uint64_t ux = wave_excl_scan(us, lane);
    for (int k = lo; k < hi; ++k) {
        const int d = ls[k];
        const uint64_t w = lu[k];
        ls[k] = dx;
        lu[k] = ux;
        dx += d;
        ux += w;
    }
    if (lane == 63) {  // the last block's running sums are the totals
        ls[L] = dx;
        lu[L] = ux;
    }
    __syncthreads();
    for (int i = lane; i < L; i += 64) {
        const int target = ls[i] + 1;
        int j = i + 1;
        while (j < L) {
            const int s = ls[j];
            if (s >= target) break;
            j += target - s;
        }
        j = min(j, L);
        const int n = j - i;
        uint64_t x = mix64((lu[j] - lu[i]) * kHashPowers.pw[i] + (uint64_t)n * 0xbf58476d1ce4e5b9ULL);
        x = x ? x : 1;
        a.hh[b + i] = x;
        a.sz[b + i] = n;
        if (n >= a.ms) a.hit[b + i] = lookup(a, x, n);
    }
}

// ---- 3. executed subtrees without a column, counted by key (one wave per
// expression) -------------------------------------------------------------------
// cnt[q] = how many spans [s + 1, s + sz[s]) of the marked positions s hold
// q, for q < L: a difference array over the spans, scanned.  Every lane must
// call it.
template <class Mark>
SR_DEV void wave_cover(const DmArgs &a, int64_t b, int L, int32_t *cnt, int32_t *part, int lane, Mark mark)
{
    for (int q = lane; q <= L; q += kWaveLanes) cnt[q] = 0;
    __syncthreads();
    for (int q = lane; q < L; q += kWaveLanes)
        if (mark(q)) {
            atomicAdd(&cnt[q + 1], 1);
            atomicSub(&cnt[min(q + a.sz[b + q], L)], 1);
        }
    __syncthreads();
    const int lo = wave_block_lo(L, lane), hi = wave_block_hi(L, lane);
    int t = 0;
    for (int q = lo; q < hi; ++q) t += cnt[q];
    int open = lane_prefix_sum(part, t, lane);
    for (int q = lo; q < hi; ++q) {
        open += cnt[q];
        cnt[q] = open;
    }
    __syncthreads();
}

// First the lookup's pending hits: the outermost ones (inside no other
// pending span) have their bodies compared by the wave; a hit inside a
// verified one is never read as a column (its positions are read through
// the outer one) and only keeps its slot alive, so it is taken as is —
// unless some outermost compare failed, then all are compared.  Then a
// position inside a hit's span is read from the column, not executed; an
// expression whose whole tree has a column and a remembered MSE runs nothing;
// the executed positions without a column are counted by key.
__global__ __launch_bounds__(64) void k_dm_count(DmArgs a)
{
    extern __shared__ int32_t lds32[];
    const int cap = a.cap;
    int32_t *cnt = lds32;  // [cap + 1]
    int32_t *lhit = cnt + cap + 1, *cp = lhit + cap, *cr = cp + cap, *cn = cr + cap, *coff = cn + cap,
            *mism = coff + cap;  // [cap] each
    __shared__ int32_t part[kWaveLanes];
    __shared__ int32_t nc, failed;
    const int e = blockIdx.x, lane = threadIdx.x;
    const int64_t b = (int64_t)e * a.cap;
    const int L = max(0, min(a.len[e], a.cap));
    if (L == 0) return;
    const uint32_t gen = (uint32_t)a.ctr[C_GEN];
    for (int q = lane; q < L; q += kWaveLanes) lhit[q] = a.hit[b + q];
    if (lane == 0) failed = 0;
    __syncthreads();
    auto pending = [&](int q) { return lhit[q] <= kHitPending; };
    auto settle = [&](int q, int slot, bool same) {
        const int h = same ? slot : -2;
        lhit[q] = h;
        a.hit[b + q] = h;
        // popular slots are hit thousands of times: store only when it changes
        if (same && a.s_seen[slot] != gen) a.s_seen[slot] = gen;
    };
    auto body2 = [&](int slot, const int32_t *&t, const float *&v) {
        t = a.s_tok + (int64_t)slot * a.cap;
        v = a.s_val + (int64_t)slot * a.cap;
    };
    // compares the pending positions that pass `pick`, settles them
    auto compare = [&](auto pick) {
        if (lane == 0) nc = 0;
        __syncthreads();
        for (int q = lane; q < L; q += kWaveLanes)
            if (pending(q) && pick(q)) {
                const int k = atomicAdd(&nc, 1);
                cp[k] = (int32_t)(b + q);
                cr[k] = kHitPending - lhit[q];
                cn[k] = a.sz[b + q];
            }
        __syncthreads();
        const int ncand = nc;
        wave_compare(a, ncand, cp, cr, cn, coff, mism, part, lane, body2);
        for (int k = lane; k < ncand; k += kWaveLanes) {
            settle((int)(cp[k] - b), cr[k], !mism[k]);
            if (mism[k]) failed = 1;
        }
        __syncthreads();
    };
    wave_cover(a, b, L, cnt, part, lane, pending);
    compare([&](int q) { return cnt[q] == 0; });
    if (failed) compare([&](int q) { return true; });
    for (int q = lane; q < L; q += kWaveLanes)
        if (pending(q)) settle(q, kHitPending - lhit[q], true);
    __syncthreads();
    const int r0 = lhit[0];
    if (r0 >= 0 && a.s_has[r0]) return;  // a settled expression runs nothing
    wave_cover(a, b, L, cnt, part, lane, [&](int q) { return lhit[q] >= 0; });
    for (int q = lane; q < L; q += kWaveLanes) {
        const int64_t p = b + q;
        if (cnt[q] > 0 || lhit[q] != -1) continue;
        const uint64_t h = a.hh[p];
        for (uint32_t i = (uint32_t)mix64(h) & (a.H2 - 1);; i = (i + 1) & (a.H2 - 1)) {
            uint64_t k = a.ct_key[i];
            if (k == 0) {
                const uint64_t prev = atomicCAS((unsigned long long *)&a.ct_key[i], 0ull, (unsigned long long)h);
                k = prev == 0 ? h : prev;
            }
            if (k != h) continue;
            // only "at least twice" is ever asked of the count, and the
            // minimum only moves down: popular keys skip most atomics
            if (a.ct_cnt[i] < 2) atomicAdd(&a.ct_cnt[i], 1);
            if ((int32_t)p < a.ct_rep[i]) atomicMin(&a.ct_rep[i], (int32_t)p);
            a.cidx[p] = (int32_t)i;
            break;
        }
    }
}

// ---- shared occurrences and their outermost shared ancestor (one wave per
// expression) -------------------------------------------------------------------
// An occurrence is "shared" when its key was executed at least twice and its
// body equals the key's first occurrence (a hash collision stays inline).
// Shared spans nest or are disjoint, so the outermost shared occurrence
// holding a shared position is the last top-level shared start at or before
// it (top-level: inside no other shared span, found by scanning a difference
// array of the spans); a top-level occurrence makes its key a root.
__global__ __launch_bounds__(64) void k_dm_share(DmArgs a)
{
    extern __shared__ int32_t lds32[];
    const int cap = a.cap;
    int32_t *code = lds32;          // [cap]: 0 not shared, 1 shared top-level, 2 shared inside
    int32_t *cnt = code + cap;      // [cap + 1]
    int32_t *cp = cnt + cap + 1, *cr = cp + cap, *cn = cr + cap, *coff = cn + cap, *mism = coff + cap;  // [cap]
    __shared__ int32_t part[kWaveLanes];
    __shared__ int32_t nc;
    const int e = blockIdx.x, lane = threadIdx.x;
    const int64_t b = (int64_t)e * a.cap;
    const int L = max(0, min(a.len[e], a.cap));
    if (L == 0) return;
    if (lane == 0) nc = 0;
    __syncthreads();
    for (int q = lane; q < L; q += kWaveLanes) {
        const int64_t p = b + q;
        const int i = a.cidx[p];
        int s = 0;
        if (i >= 0 && a.ct_cnt[i] >= 2) {
            const int64_t r = a.ct_rep[i];
            if (p == r) {
                s = 1;
            } else if (a.sz[p] == a.sz[r]) {  // body compared by the wave below
                const int k = atomicAdd(&nc, 1);
                cp[k] = (int32_t)p;
                cr[k] = (int32_t)r;
                cn[k] = a.sz[p];
            }
        }
        code[q] = s;
        cnt[q] = 0;
        if (s) a.sh[p] = 1;
    }
    if (lane == 0) cnt[L] = 0;
    __syncthreads();
    const int ncand = nc;
    wave_compare(a, ncand, cp, cr, cn, coff, mism, part, lane, [&](int r, const int32_t *&t, const float *&v) {
        t = a.tok + r;
        v = a.val + r;
    });
    for (int k = lane; k < ncand; k += kWaveLanes)
        if (!mism[k]) {
            code[cp[k] - b] = 1;
            a.sh[cp[k]] = 1;
        }
    __syncthreads();
    for (int q = lane; q < L; q += kWaveLanes)
        if (code[q]) {
            atomicAdd(&cnt[q + 1], 1);
            atomicSub(&cnt[min(q + a.sz[b + q], L)], 1);
        }
    __syncthreads();
    const int lo = wave_block_lo(L, lane), hi = wave_block_hi(L, lane);
    int t = 0;
    for (int q = lo; q < hi; ++q) t += cnt[q];
    int open = lane_prefix_sum(part, t, lane);
    int last = -1;
    for (int q = lo; q < hi; ++q) {
        open += cnt[q];
        if (code[q]) {
            code[q] = open > 0 ? 2 : 1;
            if (open == 0) last = q;
        }
    }
    // last top-level start before this lane's block (a running max)
    part[lane] = last;
    __syncthreads();
    int prev = -1;
    for (int k = 0; k < lane; ++k) prev = max(prev, part[k]);
    for (int q = lo; q < hi; ++q) {
        const int c = code[q];
        if (c == 0) continue;
        const int64_t p = b + q;
        if (c == 1) {
            prev = q;
            a.anc[p] = (int32_t)p;
            const int i = a.cidx[p];
            if ((int32_t)p < a.ct_root[i]) atomicMin(&a.ct_root[i], (int32_t)p);
        } else {
            a.anc[p] = (int32_t)(b + prev);
        }
    }
}

// ---- 4. keys, eviction ---------------------------------------------------------------
// Per key (i < H2): a root key flags its outermost first occurrence; a shared
// key with no outermost occurrence is written inside the program of the root
// holding its first occurrence, at that occurrence's offset.  Per slot
// (i < C): slots whose subtree the population no longer holds are evicted.
// The slot lookup table (i < H) is cleared for its rebuild.
__global__ __launch_bounds__(256) void k_dm_keys(DmArgs a, int32_t *blk)
{
    __shared__ int wc[4];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < a.H2 && a.ct_cnt[i] >= 2) {
        const int32_t root = a.ct_root[i];
        if (root != kNone) {
            a.flag[root] = 1;
        } else {
            const int32_t r = a.ct_rep[i];
            const int32_t A = a.anc[r];
            if (A != kNone) {
                a.ct_wr[i] = a.cidx[A];
                a.ct_off[i] = r - A;
            }
        }
    }
    bool free_slot = false;
    if (i < a.C) {
        free_slot = !a.s_live[i];
        if (!free_slot && a.s_seen[i] != (uint32_t)a.ctr[C_GEN]) {
            a.s_live[i] = 0;
            a.s_has[i] = 0;
            free_slot = true;
        }
    }
    if (i < a.H) a.ht_key[i] = 0;
    // the free-slot compaction's block counts (blocks of the slot range)
    const uint64_t m = __ballot(free_slot);
    if ((threadIdx.x & 63) == 0) wc[threadIdx.x >> 6] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0 && (int64_t)blockIdx.x * 256 < a.C) blk[blockIdx.x] = wc[0] + wc[1] + wc[2] + wc[3];
}

SR_DEV int wave_sum_i32(int v)
{
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// Ordered compaction in two passes (256 items per block): count the selected
// items of each block, then give each selected item its rank.  Selection: kind 0 = flag[i] == v over positions
// (rank[i] written), kind 1 = a free slot (s_free[rank] = i).
SR_DEV bool scan_pick(const DmArgs &a, int kind, int v, int64_t i)
{
    return kind == 0 ? a.flag[i] == v : a.s_live[i] == 0;
}

SR_DEV int64_t scan_n(const DmArgs &a, int kind) { return kind == 0 ? (int64_t)a.E * a.cap : a.C; }

__global__ __launch_bounds__(256) void k_dm_scan_count(DmArgs a, int kind, int v, int32_t *blk)
{
    __shared__ int wc[4];
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool pick = i < scan_n(a, kind) && scan_pick(a, kind, v, i);
    const uint64_t m = __ballot(pick);
    if ((threadIdx.x & 63) == 0) wc[threadIdx.x >> 6] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) blk[blockIdx.x] = wc[0] + wc[1] + wc[2] + wc[3];
}

// Each block's offset is the sum of the counts of the blocks before it
// (summed by the block itself); the last block writes the total to ctr[ci].
__global__ __launch_bounds__(256) void k_dm_scan_write(DmArgs a, int kind, int v, const int32_t *blk, int ci)
{
    __shared__ int wc[4], ws[4];
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int w = threadIdx.x >> 6;
    int before = 0;
    for (int k = threadIdx.x; k < (int)blockIdx.x; k += 256) before += blk[k];
    before = wave_sum_i32(before);
    const bool pick = i < scan_n(a, kind) && scan_pick(a, kind, v, i);
    const uint64_t m = __ballot(pick);
    if ((threadIdx.x & 63) == 0) {
        wc[w] = __popcll(m);
        ws[w] = before;
    }
    __syncthreads();
    const int base = ws[0] + ws[1] + ws[2] + ws[3];
    if (threadIdx.x == 0 && blockIdx.x == gridDim.x - 1) a.ctr[ci] = base + blk[blockIdx.x];
    if (!pick) return;
    int r = base;
    for (int k = 0; k < w; ++k) r += wc[k];
    r += __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
    if (kind == 0) a.rank[i] = r;
    else a.s_free[r] = (int32_t)i;
}

// ---- 5. slots for roots, then for nested shared keys ------------------------------
// A nested key gets a slot only when its writer root got one.
__global__ void k_dm_nested(DmArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.H2 || a.ct_cnt[i] < 2) return;
    if (a.ct_root[i] == kNone && a.ct_wr[i] >= 0 && a.ct_slot[a.ct_wr[i]] >= 0) a.flag[a.ct_rep[i]] = 2;
}

// The slot's token body is copied by k_dm_index (one wave per claimed slot).
SR_DEV void claim_slot(const DmArgs &a, int i, int slot, int64_t body)
{
    a.ct_slot[i] = slot;
    a.s_key[slot] = a.ct_key[i];
    a.s_len[slot] = a.sz[body];
    a.s_src[slot] = (int32_t)body;
    a.s_live[slot] = 1;
    a.s_has[slot] = 0;
    a.s_seen[slot] = (uint32_t)a.ctr[C_GEN];
}

__global__ void k_dm_alloc(DmArgs a, int roots)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nfree = a.ctr[C_FREE];
    const int nroot = min(min(a.ctr[C_ROOTS], nfree), a.L1);
    if (i == 0) {
        if (roots) a.ctr[C_L1] = nroot;
        else a.ctr[C_NEST] = min(a.ctr[C_NESTED], nfree - nroot);
    }
    if (i >= a.H2 || a.ct_cnt[i] < 2) return;
    if (roots) {
        if (a.ct_root[i] == kNone) return;
        const int r = a.rank[a.ct_root[i]];
        if (r >= nroot) return;
        const int slot = a.s_free[r];
        claim_slot(a, i, slot, a.ct_root[i]);
        a.ct_prog[i] = r;
        a.p1_slot[r] = slot;
        a.p1_pos[r] = a.ct_root[i];
    } else {
        if (a.flag[a.ct_rep[i]] != 2) return;
        const int r = nroot + a.rank[a.ct_rep[i]];
        if (r >= nfree) return;
        claim_slot(a, i, a.s_free[r], a.ct_rep[i]);
    }
}

// Slot lookup table rebuilt from the live slots (i < C); every shared
// position learns its key's slot (i < E * cap); the bodies of the slots
// claimed this generation (s_free[0, L1 + NEST)) are copied, a wave per slot.
__global__ void k_dm_index(DmArgs a)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    {
        const int lane = threadIdx.x & 63;
        const int claims = a.ctr[C_L1] + a.ctr[C_NEST];
        const int waves = (int)(gridDim.x * (blockDim.x / 64));
        for (int j = (int)(p / 64); j < claims; j += waves) {
            const int slot = a.s_free[j];
            const int64_t src = a.s_src[slot], dst = (int64_t)slot * a.cap;
            for (int k = lane; k < a.s_len[slot]; k += 64) {
                a.s_tok[dst + k] = a.tok[src + k];
                a.s_val[dst + k] = a.val[src + k];
            }
        }
    }
    if (p < a.C && a.s_live[p]) {
        const uint64_t h = a.s_key[p];
        for (uint32_t i = (uint32_t)mix64(h) & (a.H - 1);; i = (i + 1) & (a.H - 1)) {
            if (atomicCAS((unsigned long long *)&a.ht_key[i], 0ull, (unsigned long long)h) == 0ull) {
                a.ht_slot[i] = (int32_t)p;
                break;
            }
        }
    }
    if (p < (int64_t)a.E * a.cap && a.sh[p]) a.pslot[p] = a.ct_slot[a.cidx[p]];
}

// ---- 6. skeleton programs, emitted and compiled ------------------------------
// One wave per program (blocks < L1: launch-1 roots, then one per expression
// for launch 2).  The source span is staged in LDS; the tokens the program
// keeps follow from a scan: a position inside an OP_REF region (a hit, or
// for launch 2 also a shared subtree with a slot) is dropped, a region's
// first position becomes one OP_REF, a nested key this root writes is
// preceded by an OP_STORE, and a launch-1 program starts with the OP_STORE
// of its own column.  The skeleton is then compiled by the wave
// (sr_compile_wave.h) straight into the program buffer, its first OP_REF
// operands marked for prefetch.
__global__ __launch_bounds__(64) void k_dm_program(DmArgs a, int nv)
{
    extern __shared__ int32_t lds32[];
    const int PC = a.PC, cap = a.cap;
    int32_t *ot = lds32;
    float *ov = (float *)(ot + PC);
    int32_t *scr = (int32_t *)(ov + PC);
    const WaveCompileScratch cs = wave_compile_scratch(scr, PC);
    // emission scratch, dead before the compile reuses the area
    int32_t *xref = scr, *xst = xref + cap, *xsz = xst + cap, *xtok = xsz + cap;
    float *xval = (float *)(xtok + cap);
    int32_t *cnt = (int32_t *)(xval + cap), *pa = cnt + cap + 1, *pb = pa + kWaveLanes;
    __shared__ int total;
    DeviceWave w;
    const int g = blockIdx.x, lane = threadIdx.x;
    const bool first = g < a.L1;
    const int idx = first ? g : g - a.L1;
    int64_t P;
    int ns;
    if (first) {
        if (g >= a.ctr[C_L1]) return;
        P = a.p1_pos[g];
        ns = a.sz[P];
    } else {
        const int e = idx;
        P = (int64_t)e * cap;
        const int root = a.hit[P];
        if (root >= 0 && a.s_has[root]) {
            if (lane == 0) a.e_src[e] = -2 - root;
            return;
        }
        if (a.sh[P] && a.pslot[P] >= 0 && a.ct_prog[a.cidx[P]] >= 0) {
            if (lane == 0) {
                const int r = a.ct_prog[a.cidx[P]];
                a.e_src[e] = kFromLaunch1 | r;
                a.p1_need[r] = 1;  // this root's MSE is an expression's fitness
            }
            return;
        }
        if (lane == 0) a.e_src[e] = e;
        ns = min(a.len[e], cap);
        if (ns <= 0) return;
    }
    const int ri = first ? a.cidx[P] : 0;
    w.each([&](int l) {
        for (int x = l; x < ns; x += kWaveLanes) {
            const int64_t q = P + x;
            int ref = -1, st = -1;
            if (!first) {
                const int h = a.hit[q];
                ref = h >= 0 ? h : a.pslot[q];
            } else if (x > 0) {
                ref = a.hit[q];
                if (a.sh[q] && a.pslot[q] >= 0) {
                    const int i = a.cidx[q];
                    if (a.ct_root[i] == kNone && a.ct_wr[i] == ri && a.ct_off[i] == x) st = a.pslot[q];
                }
            }
            xref[x] = ref;
            xst[x] = st;
            xsz[x] = a.sz[q];
            xtok[x] = a.tok[q];
            xval[x] = a.val[q];
            cnt[x] = 0;
        }
        if (l == 0) cnt[ns] = 0;
    });
    // OP_REF regions: +1 after the region's first position, -1 at its end
    w.each([&](int l) {
        for (int x = l; x < ns; x += kWaveLanes)
            if (xref[x] >= 0) {
                atomicAdd(&cnt[x + 1], 1);
                atomicSub(&cnt[min(x + xsz[x], ns)], 1);
            }
    });
    w.each([&](int l) {
        int o = 0;
        for (int x = wave_block_lo(ns, l); x < wave_block_hi(ns, l); ++x) o += cnt[x];
        pa[l] = o;
    });
    // tokens kept per position (0: inside a region, 2: OP_STORE + token)
    w.each([&](int l) {
        int open = 0, c = 0;
        for (int k = 0; k < l; ++k) open += pa[k];
        for (int x = wave_block_lo(ns, l); x < wave_block_hi(ns, l); ++x) {
            open += cnt[x];
            const int e = open > 0 ? 0 : (xref[x] >= 0 ? 1 : 1 + (xst[x] >= 0));
            cnt[x] = e;
            c += e;
        }
        pb[l] = c;
    });
    w.each([&](int l) {
        int o = first ? 1 : 0;
        for (int k = 0; k < l; ++k) o += pb[k];
        for (int x = wave_block_lo(ns, l); x < wave_block_hi(ns, l); ++x) {
            const int e = cnt[x];
            if (e == 0) continue;
            if (xref[x] >= 0) {
                ot[o] = OP_REF;
                ov[o] = (float)xref[x];
                ++o;
                continue;
            }
            if (e == 2) {
                ot[o] = OP_STORE;
                ov[o] = (float)xst[x];
                ++o;
            }
            ot[o] = xtok[x];
            ov[o] = xval[x];
            ++o;
        }
        if (l == kWaveLanes - 1) total = o;
        if (first && l == 0) {
            ot[0] = OP_STORE;
            ov[0] = (float)a.p1_slot[g];
        }
    });
    const int m = total;
    int sn = 0;
    // bytecode straight to the program buffer, OP_REF prefetch list beside it
    Ins *out = (first ? a.p1_ins : a.p2_ins) + (int64_t)idx * PC;
    int32_t *pf = (first ? a.p1_pf : a.p2_pf) + (int64_t)idx * kRefPrefetch;
    int n = compile_wave(w, ot, ov, m, nv, cs, out, &sn, pf);
    if (n < 0 || sn > kStackSlots) {
        if (lane == 0) {
            atomicOr(&a.ctr[C_STATUS], 1);
            out[0].code = 0 | (SRC_CONST << 8);
            out[0].arg = f32_bits(__builtin_nanf(""));
        }
        if (lane < kRefPrefetch) pf[lane] = -1;
        n = 1;
    }
    long long refs = 0, stores = 0, ops = 0;
    for (int k = lane; k < m; k += kWaveLanes) {
        const int t = ot[k];
        refs += t == OP_REF;
        stores += t == OP_STORE;
        ops += t > 0;
    }
    refs = wave_sum(refs);
    stores = wave_sum(stores);
    ops = wave_sum(ops);
    if (lane == 0) {
        (first ? a.p1_plen : a.p2_plen)[idx] = n;
        stat_add(a.stats, g, S_EXEC, m);
        stat_add(a.stats, g, S_REFS, refs);
        stat_add(a.stats, g, S_STORES, stores);
        stat_add(a.stats, g, S_OPS, ops);
    }
}

// ---- fitness of every expression; MSEs of launch-1 roots remembered ---------------
// One wave per item.  A launch-2 program's MSE is reduced here from its
// partials, in k_finalize's order (lane-strided sums, the same shuffle tree),
// so it equals the batched evaluation's bit for bit.
__global__ __launch_bounds__(256) void k_dm_gather(DmArgs a, long long *cnt_out, const double *psum,
                                                   const long long *pcnt, int P)
{
    const int e = (int)((blockIdx.x * 256u + threadIdx.x) / 64), lane = threadIdx.x & 63;
    if (lane == 0 && e < a.L1 && e < a.ctr[C_L1] && a.p1_need[e]) {
        // only roots that were whole expressions had their MSE computed
        const int s = a.p1_slot[e];
        a.s_mse[s] = a.mse1[e];
        a.s_cnt[s] = a.cnt1[e];
        a.s_has[s] = 1;
    }
    if (e >= a.E) return;
    const int src = a.e_src[e];
    float m;
    long long c;
    if (src <= -2) {
        m = a.s_mse[-2 - src];
        c = a.s_cnt[-2 - src];
    } else if (src & kFromLaunch1) {
        m = a.mse1[src & ~kFromLaunch1];
        c = a.cnt1[src & ~kFromLaunch1];
    } else {
        double sum = 0.0;
        long long n = 0;
        for (int k = lane; k < P; k += 64) {
            sum += psum[(int64_t)e * P + k];
            n += pcnt[(int64_t)e * P + k];
        }
        sum = wave_sum(sum);
        n = wave_sum(n);
        m = n > 0 ? (float)(sum / (double)n) : qnan();
        c = n;
    }
    if (lane != 0) return;
    a.fit[e] = m;
    if (cnt_out) cnt_out[e] = c;
    stat_add(a.stats, e, S_LOGICAL, a.len[e]);
    if (e == 0) {
        a.stats[S_L1] += a.ctr[C_L1];
        a.stats[S_L2] += a.ctr[C_L2];
        a.stats[S_LIVE] = (long long)a.C - a.ctr[C_FREE] + a.ctr[C_L1] + a.ctr[C_NEST];
    }
}

// Empty slots and an empty slot lookup table (grid over max(C, H)).
__global__ void k_dm_init_slots(DmArgs a)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s == 0) {
        for (int k = 0; k < C_COUNT; ++k) a.ctr[k] = 0;
        for (int k = 0; k < S_COUNT * (1 + kStatSpread); ++k) a.stats[k] = 0;
    }
    if (s < a.H) a.ht_key[s] = 0;
    if (s >= a.C) return;
    a.s_live[s] = 0;
    a.s_has[s] = 0;
    a.s_seen[s] = 0;
    a.s_len[s] = 0;
}

unsigned grid_of(int64_t n) { return (unsigned)((n + kThreads - 1) / kThreads); }

int pow2_at_least(int64_t v)
{
    int64_t p = 64;
    while (p < v) p <<= 1;
    return (int)p;
}

DmArgs dm_args(Ctx &c, int b)
{
    DMemoState &d = c.dm;
    GaState &g = c.ga;
    DmArgs a{};
    a.tok = g.tok[b].as<int32_t>();
    a.val = g.val[b].as<float>();
    a.len = g.len[b].as<int32_t>();
    a.E = d.E;
    a.cap = d.cap;
    a.PC = d.PC;
    a.L1 = d.L1;
    a.C = d.C;
    a.H = d.H;
    a.H2 = d.H2;
    a.ms = d.min_size;
    a.sz = d.sz.as<int32_t>();
    a.hh = d.hh.as<uint64_t>();
    a.hit = d.hit.as<int32_t>();
    a.sh = d.sh.as<int32_t>();
    a.anc = d.anc.as<int32_t>();
    a.pslot = d.pslot.as<int32_t>();
    a.cidx = d.ctx_idx.as<int32_t>();
    a.flag = d.flag.as<int32_t>();
    a.rank = d.rank.as<int32_t>();
    a.s_key = d.s_key.as<uint64_t>();
    a.s_len = d.s_len.as<int32_t>();
    a.s_tok = d.s_tok.as<int32_t>();
    a.s_src = d.s_src.as<int32_t>();
    a.s_val = d.s_val.as<float>();
    a.s_seen = d.s_seen.as<uint32_t>();
    a.s_live = d.s_live.as<int32_t>();
    a.s_mse = d.s_mse.as<float>();
    a.s_cnt = d.s_cnt.as<long long>();
    a.s_has = d.s_has.as<int32_t>();
    a.s_free = d.s_free.as<int32_t>();
    a.ht_key = d.ht_key.as<uint64_t>();
    a.ht_slot = d.ht_slot.as<int32_t>();
    a.ct_key = d.ct_key.as<uint64_t>();
    a.ct_cnt = d.ct_cnt.as<int32_t>();
    a.ct_rep = d.ct_rep.as<int32_t>();
    a.ct_root = d.ct_root.as<int32_t>();
    a.ct_slot = d.ct_slot.as<int32_t>();
    a.ct_wr = d.ct_wr.as<int32_t>();
    a.ct_off = d.ct_off.as<int32_t>();
    a.ct_prog = d.ct_prog.as<int32_t>();
    a.p1_slot = d.p1_slot.as<int32_t>();
    a.p1_pos = d.p1_pos.as<int32_t>();
    a.p1_need = d.p1_need.as<int32_t>();
    a.e_src = d.e_src.as<int32_t>();
    a.p1_ins = d.p1_ins.as<Ins>();
    a.p2_ins = d.p2_ins.as<Ins>();
    a.p1_plen = d.p1_plen.as<int32_t>();
    a.p2_plen = d.p2_plen.as<int32_t>();
    a.p1_pf = d.p1_pf.as<int32_t>();
    a.p2_pf = d.p2_pf.as<int32_t>();
    a.mse1 = d.mse1.as<float>();
    a.cnt1 = d.cnt1.as<long long>();
    a.fit = c.d_mse.as<float>();
    a.ctr = d.ctr.as<int32_t>();
    a.stats = d.stats.as<long long>();
    return a;
}

}  // namespace

// Size (or re-size) the device cache for the context's GA population and
// data; columns from free HBM unless `capacity` > 0.  Not captured.
int dmemo_setup(Ctx &c, int capacity, int min_size)
{
    DMemoState &d = c.dm;
    GaState &g = c.ga;
    const int E = g.pop, cap = g.cap;
    const int PC = 2 * cap + 2;
    if (PC > MAX_EXPR_LEN)
        return c.fail(SR_ERR_UNSUPPORTED, "device-planned cache: GA stride too large for skeleton programs");
    if (d.ready && d.E == E && d.cap == cap && d.N == c.N && d.data_version == c.data_version &&
        d.min_size == min_size)
        return SR_OK;
    c.ga.release_graphs();
    int C = capacity;
    if (C <= 0) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = (size_t)1 << 34;
        free_b += d.cols.bytes;
        const double col = (double)c.N * sizeof(float);
        C = (int)std::max<int64_t>(64, std::min<int64_t>((int64_t)(0.5 * (double)free_b / col), 1 << 17));
    }
    const int64_t M = (int64_t)E * cap;
    d.E = E;
    d.cap = cap;
    d.PC = PC;
    d.L1 = 2 * E;
    d.C = C;
    d.H = pow2_at_least(2 * (int64_t)C);
    d.H2 = pow2_at_least(M);
    d.N = c.N;
    d.data_version = c.data_version;
    d.min_size = min_size;
    d.ready = false;
    d.planned = false;
    hipError_t he = hipSuccess;
    auto need = [&](DevBuf &buf, size_t bytes) {
        if (he == hipSuccess) he = buf.ensure(bytes);
    };
    d.cols.release();
    need(d.cols, (size_t)C * (size_t)c.N * 4);
    need(d.s_key, (size_t)C * 8);
    need(d.s_len, (size_t)C * 4);
    need(d.s_src, (size_t)C * 4);
    need(d.s_tok, (size_t)C * cap * 4);
    need(d.s_val, (size_t)C * cap * 4);
    need(d.s_seen, (size_t)C * 4);
    need(d.s_live, (size_t)C * 4);
    need(d.s_mse, (size_t)C * 4);
    need(d.s_cnt, (size_t)C * 8);
    need(d.s_has, (size_t)C * 4);
    need(d.s_free, (size_t)C * 4);
    need(d.ht_key, (size_t)d.H * 8);
    need(d.ht_slot, (size_t)d.H * 4);
    for (DevBuf *bb : {&d.ct_cnt, &d.ct_rep, &d.ct_root, &d.ct_slot, &d.ct_wr, &d.ct_off, &d.ct_prog})
        need(*bb, (size_t)d.H2 * 4);
    need(d.ct_key, (size_t)d.H2 * 8);
    need(d.hh, (size_t)M * 8);
    for (DevBuf *bb : {&d.sz, &d.hit, &d.sh, &d.anc, &d.pslot, &d.ctx_idx, &d.flag, &d.rank})
        need(*bb, (size_t)M * 4);
    const int64_t P1 = (int64_t)d.L1 * PC, P2 = (int64_t)E * PC;
    need(d.p1_ins, P1 * sizeof(Ins));
    need(d.p2_ins, P2 * sizeof(Ins));
    for (DevBuf *bb : {&d.p1_slot, &d.p1_pos, &d.p1_need, &d.p1_plen, &d.p1_order, &d.mse1})
        need(*bb, (size_t)d.L1 * 4);
    need(d.cnt1, (size_t)d.L1 * 8);
    need(d.p1_pf, (size_t)d.L1 * kRefPrefetch * 4);
    for (DevBuf *bb : {&d.p2_plen, &d.p2_order, &d.e_src}) need(*bb, (size_t)E * 4);
    need(d.p2_pf, (size_t)E * kRefPrefetch * 4);
    need(d.ctr, C_COUNT * 4);
    need(d.blk, (size_t)(std::max<int64_t>(M, C) / 256 + 2) * 4);
    need(d.stats, (size_t)S_COUNT * (1 + kStatSpread) * 8);
    if (he != hipSuccess) return c.hip_fail(he, "hipMalloc device cache");
    LaunchPlan lp;  // MSE partials and fitness buffers for max(launch-1 capacity, E) programs
    int rc = plan_batched(c, MODE_MSE_REF, std::max(d.L1, E), lp);
    if (rc) return rc;
    DmArgs a = dm_args(c, 0);
    hipLaunchKernelGGL(k_dm_init_slots, dim3(grid_of(std::max(C, d.H))), dim3(kThreads), 0, c.stream, a);
    if ((he = hipGetLastError()) == hipSuccess) he = hipStreamSynchronize(c.stream);
    if (he != hipSuccess) return c.hip_fail(he, "device cache init");
    d.ready = true;
    return SR_OK;
}

// Enqueue the evaluation of GA buffer b with the device cache: fitness into
// c.d_mse (count into cnt_out when set); ev (optional, 4 events) brackets the
// two interpreter launches.  Graph-capturable (without events).
int dmemo_enqueue(Ctx &c, int b, hipStream_t s, long long *cnt_out, hipEvent_t *ev)
{
    DMemoState &d = c.dm;
    if (!d.ready) return c.fail(SR_ERR_INVALID_ARG, "device cache not set up");
    const DmArgs a = dm_args(c, b);
    const int64_t M = (int64_t)d.E * d.cap;
    const int64_t big = std::max<int64_t>(M, std::max(d.L1, d.E));
    hipError_t he;
#define DM_LAUNCH(K, N, ...)                                                                     \
    do {                                                                                         \
        hipLaunchKernelGGL(K, dim3(grid_of(N)), dim3(kThreads), 0, s, a, ##__VA_ARGS__);         \
        if ((he = hipGetLastError()) != hipSuccess) return c.hip_fail(he, #K);                   \
    } while (0)
    // one wave per item with dynamic LDS
#define DM_WAVES(K, N, BYTES, ...)                                                               \
    do {                                                                                         \
        hipLaunchKernelGGL(K, dim3((unsigned)(N)), dim3(64), (BYTES), s, a, ##__VA_ARGS__);       \
        if ((he = hipGetLastError()) != hipSuccess) return c.hip_fail(he, #K);                   \
    } while (0)
    int32_t *blk = d.blk.as<int32_t>();
    // (the free-slot compaction's counts come from k_dm_keys)
    auto scan = [&](int kind, int v, int ci) -> int {
        const int64_t n = kind == 0 ? M : d.C;
        const int nb = (int)((n + 255) / 256);
        if (kind == 0) hipLaunchKernelGGL(k_dm_scan_count, dim3(nb), dim3(256), 0, s, a, kind, v, blk);
        hipLaunchKernelGGL(k_dm_scan_write, dim3(nb), dim3(256), 0, s, a, kind, v, (const int32_t *)blk, ci);
        if ((he = hipGetLastError()) != hipSuccess) return c.hip_fail(he, "device cache scan");
        return SR_OK;
    };
    const int64_t all_keys = std::max<int64_t>(d.H2, std::max(d.C, d.H));
    DM_LAUNCH(k_dm_clear, std::max<int64_t>(big, d.H2));
    DM_WAVES(k_dm_hash, d.E, 12 * ((size_t)d.cap + 1));
    DM_WAVES(k_dm_count, d.E, 4 * (7 * (size_t)d.cap + 1));
    DM_WAVES(k_dm_share, d.E, 4 * (7 * (size_t)d.cap + 1));
    DM_LAUNCH(k_dm_keys, all_keys, blk);
    if (int r = scan(1, 0, C_FREE)) return r;
    if (int r = scan(0, 1, C_ROOTS)) return r;
    DM_LAUNCH(k_dm_alloc, d.H2, 1);
    DM_LAUNCH(k_dm_nested, d.H2);
    if (int r = scan(0, 2, C_NESTED)) return r;
    DM_LAUNCH(k_dm_alloc, d.H2, 0);
    DM_LAUNCH(k_dm_index, std::max<int64_t>(d.C, M));
    // LDS: skeleton tokens and values, then the larger of the compile scratch
    // and the emission scratch that it reuses
    const int scratch = std::max(wave_compile_scratch_ints(d.PC), 6 * d.cap + 1 + 2 * kWaveLanes);
    DM_WAVES(k_dm_program, (int64_t)d.L1 + d.E, (size_t)d.PC * 8 + 4 * (size_t)scratch, c.nv);
#undef DM_WAVES
#undef DM_LAUNCH
    // both launches' program orders; their queue counters zeroed
    // (and the launch-2 program count: expressions with a program)
    if ((he = launch_order_by_len2(d.p1_plen.as<int32_t>(), d.L1, d.p1_order.as<int32_t>(),
                                   d.p2_plen.as<int32_t>(), d.E, d.p2_order.as<int32_t>(),
                                   c.d_counters.as<uint32_t>(), 2 * kQueueCounters, d.ctr.as<int32_t>() + C_L2,
                                   s)) != hipSuccess)
        return c.hip_fail(he, "device cache order");
    int P = 0;
    int rc = dmemo_launch(c, s, ev, true, &P);
    if (rc) return rc;
    const int waves = std::max(d.L1, d.E);
    hipLaunchKernelGGL(k_dm_gather, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s, a, cnt_out,
                       (const double *)c.d_psum.as<double>(), (const long long *)c.d_pcnt.as<long long>(), P);
    if ((he = hipGetLastError()) != hipSuccess) return c.hip_fail(he, "k_dm_gather");
    d.planned = true;
    return SR_OK;
}

// The two interpreter launches of the planned generation.  Running them
// again recomputes the same columns and partial sums (columns hold exactly
// what the interpreter computes, with every kernel variant), which the
// skeleton-kernel tuner relies on.
int dmemo_launch(Ctx &c, hipStream_t s, hipEvent_t *ev, bool planned, int *P_out)
{
    DMemoState &d = c.dm;
    if (!d.ready) return c.fail(SR_ERR_INVALID_ARG, "device cache not set up");
    LaunchPlan lp;
    int rc = plan_launch(c, MODE_MSE_REF, lp);
    if (rc) return rc;
    if (lp.nq > kQueueCounters) return c.fail(SR_ERR_UNSUPPORTED, "device cache: too many queues");
    if (P_out) *P_out = lp.P;
    float *cols = d.cols.as<float>();
    // launch 1: the roots (count on the device); launch 2: the expressions,
    // numbered by expression, so every expression's MSE slot is finalized
    for (int launch = 0; launch < 2; ++launch) {
        const bool one = launch == 0;
        BatchArgs ba{};
        ba.progs = one ? d.p1_ins.as<Ins>() : d.p2_ins.as<Ins>();
        ba.prog_off = nullptr;
        ba.prog_stride = d.PC;
        ba.prog_len = one ? d.p1_plen.as<int32_t>() : d.p2_plen.as<int32_t>();
        ba.order = one ? d.p1_order.as<int32_t>() : d.p2_order.as<int32_t>();
        ba.ref_pf = one ? d.p1_pf.as<int32_t>() : d.p2_pf.as<int32_t>();
        ba.cache = cols;
        ba.out = cols;
        ba.n_exprs_dev = d.ctr.as<int32_t>() + (one ? C_L1 : C_L2);
        ba.need_mse = one ? d.p1_need.as<int32_t>() : nullptr;
        // each launch its own queue counters
        EvalBufs bufs{c.d_psum.as<double>(), c.d_pcnt.as<long long>(),
                      c.d_counters.as<uint32_t>() + (one ? 0 : kQueueCounters), d.mse1.as<float>(),
                      d.cnt1.as<long long>()};
        bufs.fin_count = one ? ba.n_exprs_dev : nullptr;
        bufs.spill = c.d_spill.as<float>();
        bufs.counters_zeroed = planned;
        bufs.finalize = one;  // launch 2's partials: reduced by k_dm_gather (or unused by the tuner)
        rc = enqueue_batched(c, MODE_MSE_REF, lp, ba, one ? d.L1 : d.E, s, bufs, ev ? ev[2 * launch] : nullptr,
                             ev ? ev[2 * launch + 1] : nullptr);
        if (rc) return rc;
    }
    return SR_OK;
}

// Statistics since setup: logical / executed tokens, OP_REF / OP_STORE tokens,
// operator tokens, launch-1 / launch-2 programs, live slots, generations.
int dmemo_stats(Ctx &c, long long out[9])
{
    if (!c.dm.ready) {
        for (int k = 0; k < S_COUNT; ++k) out[k] = 0;
        return SR_OK;
    }
    long long all[S_COUNT * (1 + kStatSpread)];
    hipError_t he = hipMemcpy(all, c.dm.stats.p, sizeof(all), hipMemcpyDeviceToHost);
    if (he != hipSuccess) return c.hip_fail(he, "device cache statistics");
    for (int k = 0; k < S_COUNT; ++k) {
        out[k] = all[k];
        for (int j = 0; j < kStatSpread; ++j) out[k] += all[S_COUNT + j * S_COUNT + k];
    }
    int32_t st = 0;
    he = hipMemcpy(&st, c.dm.ctr.as<int32_t>() + C_STATUS, 4, hipMemcpyDeviceToHost);
    if (he != hipSuccess) return c.hip_fail(he, "device cache status");
    if (st) return c.fail(SR_ERR_BAD_EXPR, "device cache: a skeleton program failed to compile");
    return SR_OK;
}

}  // namespace sr
