// sr_ga.hip -- device kernels of the GA loop and the device-side program
// builder, so one generation (compile -> evaluate -> fitness -> select ->
// vary) runs without a host round trip and can be captured in a hipGraph.
#include <hip/hip_runtime.h>

#include "sr_ga.h"
#include "sr_evogp.h"
#include "sr_ga_launch.h"

namespace sr {

// One thread per individual: initial population (stream domain INIT).
__global__ void k_ga_init(GaParams p, int pop, int cap, uint64_t seed, int32_t *tok, float *val,
                          int32_t *len, int32_t *status)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= pop) return;
    Rng r;
    r.init(seed, (uint32_t)i, GA_GEN_INIT, GA_DOMAIN_INIT);
    int n = ga_generate(p, p.max_tokens, r, tok + (size_t)i * cap, val + (size_t)i * cap, cap);
    if (n < 0) {
        atomicOr((unsigned *)status, 1u);
        tok[(size_t)i * cap] = TOK_CONST;
        val[(size_t)i * cap] = 0.0f;
        n = 1;
    }
    len[i] = n;
}

// order[rank] = i under fitness_before (ascending fitness, NaN last, ties by
// index): one wave per individual, its lanes compare against n / 64 others
// each and the counts are summed across the wave.
// With a record (rec.hist set), row (*gen - gen0) of the fitness and length
// histories is written too (rows outside [gen0, gen0 + max_gens) dropped).
__global__ __launch_bounds__(256) void k_ga_rank(const float *fit, int n, int32_t *order, GaRecord rec)
{
    const int i = (int)((blockIdx.x * (unsigned)blockDim.x + threadIdx.x) / 64);
    const int lane = threadIdx.x & 63;
    if (i >= n) return;
    const float fi = fit[i];
    int rank = 0;
    for (int j = lane; j < n; j += 64) rank += fitness_before(fit[j], j, fi, i) ? 1 : 0;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) rank += __shfl_xor(rank, m, 64);
    if (lane != 0) return;
    order[rank] = i;
    if (rec.hist) {
        const uint32_t g = *rec.gen - rec.gen0;
        if (g < (uint32_t)rec.max_gens) {
            rec.hist[(size_t)g * n + i] = __float_as_uint(fi);
            rec.hist_len[(size_t)g * n + i] = (uint32_t)rec.len[i];
        }
    }
}

// One wave per child of generation *gen (per wave in dynamic LDS: the
// parents' subtree-extent prefixes, two rows of cap + 1, and the generator's
// stack).
__global__ __launch_bounds__(256) void k_ga_next(GaParams p, SelectParams s, int pop, int cap, const int32_t *tok,
                                                 const float *val, const int32_t *len, const float *fit,
                                                 const int32_t *order, uint64_t seed, const uint32_t *gen,
                                                 int32_t *otok, float *oval, int32_t *olen, int32_t *status)
{
    extern __shared__ int32_t ga_lds[];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / 64), lane = threadIdx.x & 63;
    // the child index is wave-uniform; saying so keeps the whole decision
    // path (Philox, tournaments, the generator) on the scalar unit
    const int c = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (blockDim.x / 64)) + w);
    if (c >= pop) return;
    int32_t *l1 = ga_lds + (size_t)w * ga_wave_lds_ints(cap), *l2 = l1 + cap + 1;
    int16_t *gstack = (int16_t *)(l2 + cap + 1);
    int n = ga_child_wave(p, s, pop, cap, tok, val, len, fit, order, seed, *gen, c, otok + (size_t)c * cap,
                          oval + (size_t)c * cap, l1, l2, gstack, lane);
    if (lane == 0) {
        if (n < 0) {
            atomicOr((unsigned *)status, 2u);
            otok[(size_t)c * cap] = TOK_CONST;
            oval[(size_t)c * cap] = 0.0f;
            n = 1;
        }
        olen[c] = n;
    }
}

// One thread per expression: validate and compile into the batched
// interpreter's bytecode (program e at progs + e*cap).  A malformed
// expression sets status bit 4, one needing more than 8 stack slots bit 8
// (that takes >= 2047 tokens: a need of k slots needs two subtrees of need
// k - 1 below one node, and need 0 is one operator over two leaves); either
// gets the program "LOAD NaN".
__global__ void k_compile_pop(const int32_t *tok, const float *val, const int32_t *len, int pop,
                              int cap, int nv, Ins *progs, int32_t *prog_len, int32_t *size_scratch,
                              int32_t *need_scratch, CompileFrame *frames, int32_t *max_need,
                              int32_t *status)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= pop) return;
    const int32_t *t = tok + (size_t)e * cap;
    const float *v = val + (size_t)e * cap;
    const int L = len[e];
    Ins *out = progs + (size_t)e * cap;
    int sn = 0, n = -1;
    bool ok = L > 0 && L <= cap && raw_stack_depth(t, L) >= 0;
    for (int k = 0; ok && k < L; ++k)
        if ((t[k] == TOK_VAR && !(v[k] >= 0.0f)) || t[k] == OP_REF) ok = false;
    if (ok)
        n = compile_expr(t, v, L, nv, size_scratch + (size_t)e * cap, need_scratch + (size_t)e * cap,
                         frames + (size_t)e * (cap + 1), out, &sn);
    if (n < 0 || sn > 8) {
        atomicOr((unsigned *)status, n < 0 ? 4u : 8u);
        out[0].code = 0 | (SRC_CONST << 8);
        out[0].arg = f32_bits(__builtin_nanf(""));
        n = 1;
        sn = 0;
    }
    prog_len[e] = n;
    atomicMax(max_need, sn);
}

// Longest program first (one workgroup of 1024 threads: counting sort,
// lengths < 4096).
__device__ void order_by_len_block(const int32_t *plen, int n, int32_t *order, int32_t *nonzero = nullptr)
{
    constexpr int B = 4096;
    __shared__ int hist[B];
    for (int k = threadIdx.x; k < B; k += blockDim.x) hist[k] = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        int l = plen[i];
        l = l < 0 ? 0 : (l >= B ? B - 1 : l);
        atomicAdd(&hist[B - 1 - l], 1);
    }
    __syncthreads();
    if (nonzero && threadIdx.x == 0) *nonzero = n - hist[B - 1];  // bin B - 1: length 0
    // exclusive scan of the histogram: 4 bins per thread, then the threads'
    // sums across the wave (shuffles) and across the 16 waves (LDS)
    __shared__ int wsum[1024 / 64];
    constexpr int per = B / 1024;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    int loc[per], run = 0;
#pragma unroll
    for (int k = 0; k < per; ++k) {
        loc[k] = run;
        run += hist[t * per + k];
    }
    int inc = run;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int v = __shfl_up(inc, d, 64);
        if (lane >= d) inc += v;
    }
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    int base = inc - run;
    for (int k = 0; k < w; ++k) base += wsum[k];
#pragma unroll
    for (int k = 0; k < per; ++k) hist[t * per + k] = base + loc[k];
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        int l = plen[i];
        l = l < 0 ? 0 : (l >= B ? B - 1 : l);
        order[atomicAdd(&hist[B - 1 - l], 1)] = i;
    }
}

__global__ __launch_bounds__(1024) void k_order_by_len(const int32_t *plen, int n, int32_t *order)
{
    order_by_len_block(plen, n, order);
}

// Two orders in one launch (a workgroup each); the first workgroup also
// zeroes n_zero words (the persistent launches' queue counters), the second
// writes how many of its lengths are nonzero.
__global__ __launch_bounds__(1024) void k_order_by_len2(const int32_t *plen0, int n0, int32_t *order0,
                                                        const int32_t *plen1, int n1, int32_t *order1,
                                                        uint32_t *zero, int n_zero, int32_t *nonzero1)
{
    if (blockIdx.x == 0) {
        for (int k = threadIdx.x; k < n_zero; k += blockDim.x) zero[k] = 0u;
        order_by_len_block(plen0, n0, order0);
    } else {
        order_by_len_block(plen1, n1, order1, nonzero1);
    }
}

__global__ void k_u32_add(uint32_t *p, uint32_t v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *p += v;
}

__global__ void k_u32_set(uint32_t *p, uint32_t v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

static unsigned blocks(int n, int b) { return (unsigned)((n + b - 1) / b); }

hipError_t launch_ga_init(const GaParams &p, int pop, int cap, uint64_t seed, int32_t *tok,
                          float *val, int32_t *len, int32_t *status, hipStream_t s)
{
    hipLaunchKernelGGL(k_ga_init, dim3(blocks(pop, 64)), dim3(64), 0, s, p, pop, cap, seed, tok, val,
                       len, status);
    return hipGetLastError();
}

hipError_t launch_ga_rank(const float *fit, int n, int32_t *order, hipStream_t s, const GaRecord &rec)
{
    hipLaunchKernelGGL(k_ga_rank, dim3(blocks(n, 4)), dim3(256), 0, s, fit, n, order, rec);
    return hipGetLastError();
}

hipError_t launch_ga_next(const GaParams &p, const SelectParams &sp, int pop, int cap,
                          const int32_t *tok, const float *val, const int32_t *len, const float *fit,
                          const int32_t *order, uint64_t seed, const uint32_t *gen, int32_t *otok,
                          float *oval, int32_t *olen, int32_t *status, hipStream_t s)
{
    hipLaunchKernelGGL(k_ga_next, dim3(blocks(pop, 4)), dim3(256), (size_t)4 * ga_wave_lds_ints(cap) * sizeof(int32_t), s, p,
                       sp, pop, cap, tok, val, len, fit, order, seed, gen, otok, oval, olen, status);
    return hipGetLastError();
}

hipError_t launch_compile_pop(const int32_t *tok, const float *val, const int32_t *len, int pop,
                              int cap, int nv, Ins *progs, int32_t *prog_len, int32_t *size_scratch,
                              int32_t *need_scratch, CompileFrame *frames, int32_t *max_need,
                              int32_t *status, hipStream_t s)
{
    hipLaunchKernelGGL(k_compile_pop, dim3(blocks(pop, 64)), dim3(64), 0, s, tok, val, len, pop, cap,
                       nv, progs, prog_len, size_scratch, need_scratch, frames, max_need, status);
    return hipGetLastError();
}

hipError_t launch_order_by_len(const int32_t *plen, int n, int32_t *order, hipStream_t s)
{
    hipLaunchKernelGGL(k_order_by_len, dim3(1), dim3(1024), 0, s, plen, n, order);
    return hipGetLastError();
}

hipError_t launch_order_by_len2(const int32_t *plen0, int n0, int32_t *order0, const int32_t *plen1, int n1,
                                int32_t *order1, uint32_t *zero, int n_zero, int32_t *nonzero1, hipStream_t s)
{
    hipLaunchKernelGGL(k_order_by_len2, dim3(2), dim3(1024), 0, s, plen0, n0, order0, plen1, n1, order1, zero,
                       n_zero, nonzero1);
    return hipGetLastError();
}

hipError_t launch_u32_add(uint32_t *p, uint32_t v, hipStream_t s)
{
    hipLaunchKernelGGL(k_u32_add, dim3(1), dim3(64), 0, s, p, v);
    return hipGetLastError();
}

hipError_t launch_u32_set(uint32_t *p, uint32_t v, hipStream_t s)
{
    hipLaunchKernelGGL(k_u32_set, dim3(1), dim3(64), 0, s, p, v);
    return hipGetLastError();
}

}  // namespace sr

namespace sr {

// Batch subtree replacement (the evogp mutate / crossover kernels of
// mutation.cu:75-125 and crossover.cu:42-69 restated on prefix tokens):
// child n = left[ln] with the subtree at node lidx[n] replaced by the subtree
// of right[rn] at node ridx[n].  A child that would exceed cap, or an index
// outside its tree, copies the left tree unchanged.
__global__ void k_splice(int n, int cap, const int32_t *ltok, const float *lval, const int32_t *llen,
                         const int32_t *lsel, const int32_t *lidx, const int32_t *rtok,
                         const float *rval, const int32_t *rlen, const int32_t *rsel,
                         const int32_t *ridx, int32_t *otok, float *oval, int32_t *olen)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int32_t *lt = ltok + (size_t)lsel[k] * cap;
    const float *lv = lval + (size_t)lsel[k] * cap;
    const int L = llen[lsel[k]];
    const int32_t *rt = rtok + (size_t)rsel[k] * cap;
    const float *rv = rval + (size_t)rsel[k] * cap;
    const int R = rlen[rsel[k]];
    int32_t *ct = otok + (size_t)k * cap;
    float *cv = oval + (size_t)k * cap;
    const int li = lidx[k], ri = ridx[k];
    int out;
    if (li < 0 || li >= L || ri < 0 || ri >= R) {
        out = ga_copy(lt, lv, L, ct, cv);
    } else {
        const int ls = ga_subtree_size(lt, L, li);
        const int rs = ga_subtree_size(rt, R, ri);
        const int total = L - ls + rs;
        if (total > cap) {
            out = ga_copy(lt, lv, L, ct, cv);
        } else {
            ga_copy(lt, lv, li, ct, cv);
            ga_copy(rt + ri, rv + ri, rs, ct + li, cv + li);
            ga_copy(lt + li + ls, lv + li + ls, L - li - ls, ct + li + rs, cv + li + rs);
            out = total;
        }
    }
    olen[k] = out;
}

hipError_t launch_splice(int n, int cap, const int32_t *ltok, const float *lval, const int32_t *llen,
                         const int32_t *lsel, const int32_t *lidx, const int32_t *rtok,
                         const float *rval, const int32_t *rlen, const int32_t *rsel,
                         const int32_t *ridx, int32_t *otok, float *oval, int32_t *olen, hipStream_t s)
{
    hipLaunchKernelGGL(k_splice, dim3(blocks(n, 64)), dim3(64), 0, s, n, cap, ltok, lval, llen, lsel,
                       lidx, rtok, rval, rlen, rsel, ridx, otok, oval, olen);
    return hipGetLastError();
}

// evogp layout <-> prefix streams, one thread per individual.
__global__ void k_evogp_to_prefix(int pop, int max_len, const float *value, const int16_t *type,
                                  const int16_t *size, int cap, int32_t *tok, float *val, int32_t *len)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= pop) return;
    const size_t a = (size_t)i * max_len, b = (size_t)i * cap;
    len[i] = evogp_to_prefix_one(value + a, type + a, size + a, max_len, cap, tok + b, val + b);
}

__global__ void k_prefix_to_evogp(int pop, int cap, const int32_t *tok, const float *val, const int32_t *len,
                                  int max_len, float *value, int16_t *type, int16_t *size)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= pop) return;
    const size_t a = (size_t)i * cap, b = (size_t)i * max_len;
    const int n = len[i] < cap ? len[i] : cap;
    prefix_to_evogp_one(tok + a, val + a, n, max_len, value + b, type + b, size + b);
}

hipError_t launch_evogp_to_prefix(int pop, int max_len, const float *value, const int16_t *type,
                                  const int16_t *size, int cap, int32_t *tok, float *val, int32_t *len,
                                  hipStream_t s)
{
    hipLaunchKernelGGL(k_evogp_to_prefix, dim3(blocks(pop, 64)), dim3(64), 0, s, pop, max_len, value, type,
                       size, cap, tok, val, len);
    return hipGetLastError();
}

hipError_t launch_prefix_to_evogp(int pop, int cap, const int32_t *tok, const float *val, const int32_t *len,
                                  int max_len, float *value, int16_t *type, int16_t *size, hipStream_t s)
{
    hipLaunchKernelGGL(k_prefix_to_evogp, dim3(blocks(pop, 64)), dim3(64), 0, s, pop, cap, tok, val, len,
                       max_len, value, type, size);
    return hipGetLastError();
}

}  // namespace sr
