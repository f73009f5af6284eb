// sr_cache.cpp -- cross-population subtree cache (the reference's
// evaluate_gpu_mse, cl_gpu_eval.cu:332-564, and detect_common_subtrees,
// detect.cpp:120-387).
//
// Per evaluation:
//   1. detect common subtrees (size >= min_size, >= min_freq exact
//      occurrences) with Merkle hashes over (token, double value);
//   2. give every detected subtree not yet cached a slot, keyed by its flat
//      FNV hash (cl_gpu_eval.cu:384-397), while slots remain; unlike the
//      reference, a slot also holds the subtree's tokens and values and a
//      lookup compares them, because the flat hash collides on subtrees whose
//      values differ only in high bits (variables 0 and 2: 0.0 and 2.0);
//   3. mark, left to right, the first cached subtree of size >= min_size at
//      each position (cl_gpu_eval.cu:294-329) and build skeletons in which a
//      cached subtree is one OP_REF token (cl_gpu_eval.cu:433-459);
//   4. GPU: evaluate the new subtrees into their value columns [slot][N],
//      then the skeletons with OP_REF reading those columns, fused MSE.
// The cache persists across calls (generations) until sr_cache_clear or new
// data.  Columns are evaluated by the same interpreter on the same operands,
// so cached results equal uncached ones bit for bit.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <numeric>

#include "sr_ctx.h"
#include "sr_kernels.h"
#include "sr_launch.h"

namespace sr {

static constexpr uint64_t kFnvOffset = 14695981039346656037ULL;
static constexpr uint64_t kFnvPrime = 1099511628211ULL;

static inline uint64_t dbits(double d)
{
    uint64_t u;
    std::memcpy(&u, &d, 8);
    return u;
}

// FNV over (token widened as signed, double value bits) pairs.
static uint64_t flat_hash(const int32_t *t, const double *v, int n)
{
    uint64_t h = kFnvOffset;
    for (int i = 0; i < n; ++i) {
        h ^= (uint64_t)(int64_t)t[i];
        h *= kFnvPrime;
        h ^= dbits(v[i]);
        h *= kFnvPrime;
    }
    return h;
}

// Detection arity: tokens <= 0 are leaves, < 10 binary, else unary.
static inline int det_arity(int t) { return t <= 0 ? 0 : (t < 10 ? 2 : 1); }

// Sizes (truncated at the stream end like detect.cpp:23-47) and Merkle
// hashes (detect.cpp:97-118) of every position.
static void merkle(const int32_t *t, const double *v, int n, uint64_t *h, int32_t *sz)
{
    for (int i = n - 1; i >= 0; --i) {
        uint64_t x = kFnvOffset;
        x ^= (uint64_t)(int64_t)t[i];
        x *= kFnvPrime;
        x ^= dbits(v[i]);
        x *= kFnvPrime;
        int s = 1, pos = i + 1;
        for (int k = 0; k < det_arity(t[i]); ++k) {
            if (pos >= n) break;
            x = x ^ (h[pos] + 0x9e3779b97f4a7c15ULL + (x << 6) + (x >> 2));
            s += sz[pos];
            pos += sz[pos];
        }
        h[i] = x;
        sz[i] = s;
    }
}

struct Detected {
    int expr, start, size;
    uint64_t merkle, flat;
    int count;
};

struct HostPop {
    const int32_t *tok;
    std::vector<double> val;  // packed double values
    const int32_t *off;
    const int32_t *len;
    int E;
    int64_t T;    // total tokens
    int64_t end;  // packed array extent (max offset + length)
};

static void host_pop(const sr_population &pop, HostPop &hp)
{
    hp.tok = pop.tokens;
    hp.off = pop.offsets;
    hp.len = pop.lengths;
    hp.E = pop.num_exprs;
    hp.T = 0;
    int64_t end = 0;
    for (int e = 0; e < hp.E; ++e) {
        hp.T += pop.lengths[e];
        end = std::max<int64_t>(end, (int64_t)pop.offsets[e] + pop.lengths[e]);
    }
    hp.end = end;
    hp.val.resize(end);
    if (pop.values) std::memcpy(hp.val.data(), pop.values, end * 8);
    else
        for (int64_t i = 0; i < end; ++i) hp.val[i] = (double)pop.values_f32[i];
}

// detect_common_subtrees: groups of >= min_freq verified-equal subtrees,
// ordered by size descending (ties: Merkle hash ascending).  hints (packed
// like the population) receive the greedy non-overlapping marking.
static void detect(const HostPop &p, int min_size, int min_freq, std::vector<Detected> &out,
                   std::vector<int32_t> *hints, int64_t *covered)
{
    struct Cand {
        uint64_t h;
        int32_t expr, start, size;
    };
    std::vector<Cand> cand;
    cand.reserve(p.T);
    std::vector<uint64_t> hs;
    std::vector<int32_t> sz;
    for (int e = 0; e < p.E; ++e) {
        const int n = p.len[e];
        hs.resize(n);
        sz.resize(n);
        merkle(p.tok + p.off[e], p.val.data() + p.off[e], n, hs.data(), sz.data());
        for (int i = n - 1; i >= 0; --i)
            if (sz[i] >= min_size) cand.push_back({hs[i], e, i, sz[i]});
    }
    std::sort(cand.begin(), cand.end(), [](const Cand &a, const Cand &b) {
        if (a.h != b.h) return a.h < b.h;
        if (a.expr != b.expr) return a.expr < b.expr;
        return a.start < b.start;
    });
    out.clear();
    std::vector<std::pair<int, int>> occ;  // (expr, start) per verified group, flattened
    std::vector<int> group_first;
    size_t i = 0;
    while (i < cand.size()) {
        size_t j = i + 1;
        while (j < cand.size() && cand[j].h == cand[i].h) ++j;
        if ((int)(j - i) >= min_freq) {
            const Cand &r = cand[i];
            const int32_t *rt = p.tok + p.off[r.expr] + r.start;
            const double *rv = p.val.data() + p.off[r.expr] + r.start;
            const size_t first = occ.size();
            for (size_t k = i; k < j; ++k) {
                const Cand &c = cand[k];
                if (c.size != r.size) continue;
                const int32_t *ct = p.tok + p.off[c.expr] + c.start;
                const double *cv = p.val.data() + p.off[c.expr] + c.start;
                if (std::memcmp(ct, rt, (size_t)r.size * 4) != 0) continue;
                bool same = true;
                for (int q = 0; q < r.size && same; ++q) same = cv[q] == rv[q];
                if (same) occ.push_back({c.expr, c.start});
            }
            if ((int)(occ.size() - first) >= min_freq) {
                Detected d;
                d.expr = r.expr;
                d.start = r.start;
                d.size = r.size;
                d.merkle = r.h;
                d.flat = flat_hash(rt, rv, r.size);
                d.count = (int)(occ.size() - first);
                out.push_back(d);
                group_first.push_back((int)first);
            } else {
                occ.resize(first);
            }
        }
        i = j;
    }
    std::vector<int> idx(out.size());
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&](int a, int b) {
        if (out[a].size != out[b].size) return out[a].size > out[b].size;
        return out[a].merkle < out[b].merkle;
    });
    std::vector<Detected> sorted;
    std::vector<int> sorted_first;
    for (int k : idx) {
        sorted.push_back(out[k]);
        sorted_first.push_back(group_first[k]);
    }
    out.swap(sorted);
    if (hints) {
        hints->assign((size_t)p.end, -1);
        for (size_t s = 0; s < out.size(); ++s) {
            for (int k = 0; k < out[s].count; ++k) {
                auto [e, st] = occ[sorted_first[s] + k];
                int32_t *h = hints->data() + p.off[e] + st;
                bool overlap = false;
                for (int q = 0; q < out[s].size && !overlap; ++q) overlap = h[q] != -1;
                if (overlap) continue;
                h[0] = (int32_t)s;
                for (int q = 1; q < out[s].size; ++q) h[q] = -2;
            }
        }
        if (covered) {
            int64_t c = 0;
            for (int e = 0; e < p.E; ++e)
                for (int q = 0; q < p.len[e]; ++q) c += (*hints)[p.off[e] + q] != -1;
            *covered = c;
        }
    }
}

// Slot holding exactly this subtree, or -1.
static int find_slot(const SubtreeCache &cache, uint64_t h, const int32_t *t, const double *v, int size)
{
    auto r = cache.slot_of.equal_range(h);
    for (auto it = r.first; it != r.second; ++it) {
        const int s = it->second;
        if (cache.body_len[s] == size &&
            std::memcmp(cache.body_tok.data() + cache.body_off[s], t, sizeof(int32_t) * size) == 0 &&
            std::memcmp(cache.body_val.data() + cache.body_off[s], v, sizeof(double) * size) == 0)
            return s;
    }
    return -1;
}

// mark_cached_subtrees_from_state: hints[pos] = slot / -2 / -1.
static void mark_cached(const HostPop &p, const SubtreeCache &cache, int min_size,
                        std::vector<int32_t> &hints)
{
    hints.assign((size_t)p.end, -1);
    std::vector<uint64_t> hs;
    std::vector<int32_t> sz;
    for (int e = 0; e < p.E; ++e) {
        const int n = p.len[e];
        const int32_t *t = p.tok + p.off[e];
        const double *v = p.val.data() + p.off[e];
        int32_t *h = hints.data() + p.off[e];
        hs.resize(n);
        sz.resize(n);
        merkle(t, v, n, hs.data(), sz.data());
        for (int pos = 0; pos < n; ++pos) {
            if (h[pos] != -1) continue;
            const int size = sz[pos];
            if (size < min_size) continue;
            if (pos + size > n) continue;
            const int slot = find_slot(cache, flat_hash(t + pos, v + pos, size), t + pos, v + pos, size);
            if (slot < 0) continue;
            h[pos] = slot;
            for (int k = 1; k < size && pos + k < n; ++k) h[pos + k] = -2;
        }
    }
}

static double ms_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

int eval_cached(Ctx &c, const sr_population &pop, sr_stats *st)
{
    SubtreeCache &cache = c.cache;
    if (cache.data_version != c.data_version || cache.N != c.N) {
        cache.clear();
        cache.data_version = c.data_version;
        cache.N = c.N;
    }
    auto t0 = std::chrono::steady_clock::now();
    HostPop hp;
    host_pop(pop, hp);
    // Validate the token streams before hashing them.
    Program check;
    int rc = build_program(c, pop, true, check);
    if (rc) return rc;

    std::vector<Detected> det;
    int64_t covered = 0;
    std::vector<int32_t> det_hints;
    detect(hp, cache.min_size, cache.min_freq, det, &det_hints, &covered);

    // New slots for detected subtrees not cached yet.
    std::vector<int> new_sub, new_slot;
    for (size_t s = 0; s < det.size(); ++s) {
        const int32_t *bt = hp.tok + hp.off[det[s].expr] + det[s].start;
        const double *bv = hp.val.data() + hp.off[det[s].expr] + det[s].start;
        if (find_slot(cache, det[s].flat, bt, bv, det[s].size) >= 0) continue;
        if ((int)cache.keys.size() >= cache.capacity) continue;
        int slot = (int)cache.keys.size();
        cache.keys.push_back(det[s].flat);
        cache.slot_of.emplace(det[s].flat, slot);
        cache.body_off.push_back((int32_t)cache.body_tok.size());
        cache.body_len.push_back(det[s].size);
        cache.body_tok.insert(cache.body_tok.end(), bt, bt + det[s].size);
        cache.body_val.insert(cache.body_val.end(), bv, bv + det[s].size);
        new_sub.push_back((int)s);
        new_slot.push_back(slot);
    }
    mark_cached(hp, cache, cache.min_size, cache.last_hints);

    // Skeleton population: cached subtree -> (OP_REF, slot).
    std::vector<int32_t> sk_tok, sk_off(hp.E), sk_len(hp.E);
    std::vector<double> sk_val;
    sk_tok.reserve(hp.T);
    sk_val.reserve(hp.T);
    for (int e = 0; e < hp.E; ++e) {
        sk_off[e] = (int32_t)sk_tok.size();
        const int32_t *t = hp.tok + hp.off[e];
        const double *v = hp.val.data() + hp.off[e];
        const int32_t *h = cache.last_hints.data() + hp.off[e];
        for (int k = 0; k < hp.len[e]; ++k) {
            if (h[k] >= 0) {
                sk_tok.push_back(OP_REF);
                sk_val.push_back((double)h[k]);
                int size = 1, need = 1, q = k;
                while (need > 0 && q < hp.len[e]) {
                    need += det_arity(t[q]) - 1;
                    ++q;
                }
                size = q - k;
                k += size - 1;
            } else if (h[k] == -1) {
                sk_tok.push_back(t[k]);
                sk_val.push_back(v[k]);
            }
        }
        sk_len[e] = (int32_t)sk_tok.size() - sk_off[e];
    }
    // New-subtree population.
    std::vector<int32_t> nt, no(new_sub.size()), nl(new_sub.size());
    std::vector<double> nvv;
    for (size_t k = 0; k < new_sub.size(); ++k) {
        const Detected &d = det[new_sub[k]];
        no[k] = (int32_t)nt.size();
        nl[k] = d.size;
        const int32_t *t = hp.tok + hp.off[d.expr] + d.start;
        const double *v = hp.val.data() + hp.off[d.expr] + d.start;
        nt.insert(nt.end(), t, t + d.size);
        nvv.insert(nvv.end(), v, v + d.size);
    }
    const double detect_ms = ms_since(t0);

    hipError_t he = cache.cols.ensure((size_t)cache.capacity * (size_t)c.N * sizeof(float));
    if (he != hipSuccess) return c.hip_fail(he, "hipMalloc cache columns");

    float kernel_ms = 0.f, xfer_ms = 0.f, ms = 0.f;
    if (!new_sub.empty()) {
        sr_population np{(int32_t)new_sub.size(), nt.data(), nvv.data(), nullptr, no.data(), nl.data()};
        Program pn;
        rc = build_program(c, np, false, pn);
        if (rc) return rc;
        rc = run_batched(c, pn, MODE_COLUMN, nullptr, cache.cols.as<float>(), new_slot.data(), st);
        if (rc) return rc;
        if ((he = hipStreamSynchronize(c.stream)) != hipSuccess) return c.hip_fail(he, "subtree columns");
        hipEventElapsedTime(&ms, c.ev[0], c.ev[1]);
        xfer_ms += ms;
        hipEventElapsedTime(&ms, c.ev[1], c.ev[2]);
        kernel_ms += ms;
    }
    sr_population sp{hp.E, sk_tok.data(), sk_val.data(), nullptr, sk_off.data(), sk_len.data()};
    Program ps;
    rc = build_program(c, sp, false, ps, (int)cache.keys.size());
    if (rc) return rc;
    rc = run_batched(c, ps, MODE_MSE_REF, cache.cols.as<float>(), nullptr, nullptr, st);
    if (rc) return rc;
    if (st) {
        st->detect_time_ms += detect_ms;
        st->transfer_time_ms += xfer_ms;
        st->kernel_time_ms += kernel_ms;  // sr_eval_mse adds the skeleton pass
        st->num_subtrees = (int32_t)det.size();
        double sum = 0;
        for (auto &d : det) sum += d.size;
        st->avg_subtree_size = det.empty() ? 0.0 : sum / det.size();
        st->coverage = hp.T > 0 ? (double)covered / (double)hp.T : 0.0;
        st->num_new_subtrees = (int32_t)new_sub.size();
        st->num_cached = (int32_t)cache.keys.size();
    }
    return SR_OK;
}

// Stateless detection for inspection / parity checks.
int detect_population(Ctx &c, const sr_population &pop, int min_size, int min_freq, int32_t *hints_out,
                      uint64_t *flat_out, int32_t *size_out, int32_t *count_out, int max_subs,
                      int32_t *num_subs)
{
    Program check;
    int rc = build_program(c, pop, true, check);
    if (rc) return rc;
    HostPop hp;
    host_pop(pop, hp);
    std::vector<Detected> det;
    std::vector<int32_t> hints;
    detect(hp, min_size, min_freq, det, &hints, nullptr);
    if (num_subs) *num_subs = (int32_t)det.size();
    for (int s = 0; s < (int)det.size() && s < max_subs; ++s) {
        if (flat_out) flat_out[s] = det[s].flat;
        if (size_out) size_out[s] = det[s].size;
        if (count_out) count_out[s] = det[s].count;
    }
    if (hints_out) std::memcpy(hints_out, hints.data(), hints.size() * 4);
    return SR_OK;
}

}  // namespace sr

using namespace sr;

extern "C" {

int sr_cache_config(sr_ctx *x, int32_t capacity, int32_t min_size, int32_t min_freq)
{
    if (!x || capacity < 1 || min_size < 1 || min_freq < 1) return SR_ERR_INVALID_ARG;
    Ctx &c = x->c;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    if (capacity != c.cache.capacity) {
        c.cache.cols.release();
        c.cache.clear();
    }
    c.cache.capacity = capacity;
    c.cache.min_size = min_size;
    c.cache.min_freq = min_freq;
    return SR_OK;
}

int sr_cache_clear(sr_ctx *x)
{
    if (!x) return SR_ERR_INVALID_ARG;
    std::lock_guard<std::recursive_mutex> lk(x->c.mu);
    x->c.cache.clear();
    x->c.memo.planner.forget();
    return SR_OK;
}

int sr_cache_state(sr_ctx *x, uint64_t *keys_out, int32_t max_keys, int32_t *num_cached)
{
    if (!x) return SR_ERR_INVALID_ARG;
    Ctx &c = x->c;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    if (num_cached) *num_cached = (int32_t)c.cache.keys.size();
    if (keys_out)
        for (int i = 0; i < (int)c.cache.keys.size() && i < max_keys; ++i) keys_out[i] = c.cache.keys[i];
    return SR_OK;
}

int sr_cache_last_hints(sr_ctx *x, int32_t *hints_out, int64_t max_len, int64_t *len_out)
{
    if (!x) return SR_ERR_INVALID_ARG;
    Ctx &c = x->c;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    const auto &h = c.cache.last_hints;
    if (len_out) *len_out = (int64_t)h.size();
    if (hints_out) std::memcpy(hints_out, h.data(), std::min<int64_t>(max_len, (int64_t)h.size()) * 4);
    return SR_OK;
}

int sr_detect(sr_ctx *x, const sr_population *pop, int32_t min_size, int32_t min_freq, int32_t *hints_out,
              uint64_t *flat_out, int32_t *size_out, int32_t *count_out, int32_t max_subs, int32_t *num_subs)
{
    if (!x || !pop || min_size < 1 || min_freq < 1) return SR_ERR_INVALID_ARG;
    Ctx &c = x->c;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    return detect_population(c, *pop, min_size, min_freq, hints_out, flat_out, size_out, count_out, max_subs,
                             num_subs);
}

}  // extern "C"
