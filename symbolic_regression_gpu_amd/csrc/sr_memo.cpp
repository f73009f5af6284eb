// sr_memo.cpp -- generational subtree cache: the host planner (sr_memo.h),
// its evaluation over device columns, and the C ABI (sr_cache_policy, the
// sr_memo_planner_* inspection entry points).
//
// No counterpart in the reference: its cache (cl_gpu_eval.cu:294-329,
// 384-397) keeps the first subtrees it detects and never evicts.  This one
// follows the population: between GA generations most subtrees survive, so a
// generation executes the changed nodes and reads the rest from columns.
// Columns hold exactly what the interpreter computes for a subtree, so every
// MSE equals the uncached batched one bit for bit.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <thread>

#include "sr_ctx.h"
#include "sr_kernels.h"
#include "sr_launch.h"
#include "sr_memo.h"

namespace sr {

// ---- FlatMap ----------------------------------------------------------------
void FlatMap::reset(size_t expected)
{
    size_t cap = 64;
    while (cap < 2 * expected) cap <<= 1;
    cap = std::max(cap, keys_.size());  // never shrinks: no reallocation per call
    if (keys_.size() != cap) {
        keys_.assign(cap, 0);
        vals_.assign(cap, kEmpty);
    } else {
        std::fill(vals_.begin(), vals_.end(), kEmpty);
    }
    mask_ = cap - 1;
    n_ = 0;
}

static inline uint64_t fmix64(uint64_t x)
{
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33;
    return x;
}

int32_t FlatMap::find(uint64_t k) const
{
    if (keys_.empty()) return -1;
    for (size_t i = fmix64(k) & mask_;; i = (i + 1) & mask_) {
        if (vals_[i] == kEmpty) return -1;
        if (keys_[i] == k) return vals_[i];
    }
}

void FlatMap::grow()
{
    std::vector<uint64_t> k = std::move(keys_);
    std::vector<int32_t> v = std::move(vals_);
    keys_.assign(k.size() * 2, 0);
    vals_.assign(k.size() * 2, kEmpty);
    mask_ = keys_.size() - 1;
    n_ = 0;
    for (size_t i = 0; i < k.size(); ++i)
        if (v[i] != kEmpty) put(k[i], v[i]);
}

void FlatMap::put(uint64_t k, int32_t v)
{
    if (keys_.empty()) reset(16);
    if (2 * (n_ + 1) > keys_.size()) grow();
    for (size_t i = fmix64(k) & mask_;; i = (i + 1) & mask_) {
        if (vals_[i] == kEmpty) {
            keys_[i] = k;
            vals_[i] = v;
            ++n_;
            return;
        }
        if (keys_[i] == k) {
            vals_[i] = v;
            return;
        }
    }
}

// ---- hashing ----------------------------------------------------------------
static inline uint64_t dbits(double d)
{
    uint64_t u;
    std::memcpy(&u, &d, 8);
    return u;
}

// Tokens <= 0 are leaves, 1..9 binary, the rest unary (the evaluator's
// arity, cl_cpu_eval.cpp:82).
static inline int arity(int t) { return t <= 0 ? 0 : (t < 10 ? 2 : 1); }

// Subtree sizes and structural hashes of every position: a non-commutative
// chain over (token, value bits, child hashes) through a full 64-bit
// finaliser.  The reference's Merkle combine (detect.cpp:97-118) maps small
// mirrored trees such as sub(x2, mul(x0, x2)) and sub(x0, mul(x2, x0)) to one
// hash, so it is not used as a cache key.
static void strong_hash(const int32_t *t, const double *v, int n, uint64_t *h, int32_t *sz)
{
    for (int i = n - 1; i >= 0; --i) {
        uint64_t x = fmix64((uint64_t)(uint32_t)t[i] * 0x9e3779b97f4a7c15ULL ^
                            fmix64(dbits(v[i]) + 0x632be59bd9b4e019ULL));
        int s = 1, pos = i + 1;
        for (int k = 0; k < arity(t[i]); ++k) {
            if (pos >= n) break;
            x = fmix64(x * 0xbf58476d1ce4e5b9ULL + h[pos] + (uint64_t)(k + 1));
            s += sz[pos];
            pos += sz[pos];
        }
        h[i] = x;
        sz[i] = s;
    }
}

// ---- planner ------------------------------------------------------------------
void MemoPlanner::configure(int capacity, int min_size, int store_policy)
{
    cap_ = capacity;
    min_size_ = min_size;
    policy_ = store_policy;
    forget();
}

void MemoPlanner::forget()
{
    slot_of_.reset(0);
    prev_bits_.clear();
    key_of_.clear();
    born_epoch_.clear();
    seen_gen_.clear();
    live_.clear();
    has_mse_.clear();
    body_tok_.clear();
    body_val_.clear();
    mse_of_.clear();
    cnt_of_.clear();
    free_.clear();
    used_ = 0;
    live_count_ = 0;
}

int MemoPlanner::alloc(uint64_t key, const int32_t *tok, const double *val, int n)
{
    int s;
    if (!free_.empty()) {
        s = free_.back();
        free_.pop_back();
    } else if (used_ < cap_) {
        s = used_++;
        key_of_.push_back(0);
        born_epoch_.push_back(0);
        seen_gen_.push_back(0);
        live_.push_back(0);
        has_mse_.push_back(0);
        body_tok_.emplace_back();
        body_val_.emplace_back();
        mse_of_.push_back(0.f);
        cnt_of_.push_back(0);
    } else {
        return -1;
    }
    key_of_[s] = key;
    born_epoch_[s] = epoch_;
    seen_gen_[s] = gen_;
    live_[s] = 1;
    has_mse_[s] = 0;
    body_tok_[s].assign(tok, tok + n);
    body_val_[s].assign(val, val + n);
    slot_of_.put(key, s);
    ++live_count_;
    return s;
}

void MemoPlanner::plan(const sr_population &pop, MemoPlan &out, const std::function<void(int)> &launch_ready)
{
    const int E = pop.num_exprs;
    const int ms = min_size_;
    out.tok.clear();
    out.val.clear();
    out.off.clear();
    out.len.clear();
    out.root_slot.clear();
    out.launch_begin.assign(1, 0);
    out.src.assign(E, -1);
    out.stores = 0;

    int64_t end = 0, T = 0;
    for (int e = 0; e < E; ++e) {
        end = std::max<int64_t>(end, (int64_t)pop.offsets[e] + pop.lengths[e]);
        T += pop.lengths[e];
    }
    const int32_t *tok = pop.tokens;
    const double *vals = pop.values;
    if (!vals) {
        vals_.resize(end);
        for (int64_t i = 0; i < end; ++i) vals_[i] = (double)pop.values_f32[i];
        vals = vals_.data();
    }
    h_.resize(end);
    sz_.resize(end);
    uint64_t *h = h_.data();
    int32_t *sz = sz_.data();

    // >= 0: slot with a matching body; -1: no slot; -2: the key belongs to
    // another body (that subtree is never cached).  Reads planner state only.
    auto lookup = [&](int64_t q) -> int {
        const int s = slot_of_.find(h[q]);
        if (s < 0) return -1;
        const int n = sz[q];
        if ((int)body_tok_[s].size() != n) return -2;
        if (std::memcmp(body_tok_[s].data(), tok + q, (size_t)n * 4) != 0) return -2;
        const double *bv = body_val_[s].data();
        for (int k = 0; k < n; ++k)
            if (dbits(bv[k]) != dbits(vals[q + k])) return -2;
        return s;
    };

    // Hash every position and mark the slots this population uses, settle
    // expressions by remembered MSEs and list the positions the others will
    // execute without a column (threads over contiguous expression ranges;
    // planner state is only read), then evict the slots not met.  A slot met
    // here survives the eviction, so these lookups stay valid after it.
    ++gen_;
    const bool reused = policy_ == SR_CACHE_GENERATIONAL_REUSED || policy_ == SR_CACHE_GENERATIONAL_SHARED;
    // keys met in this population, as a 2^21-bit filter (a false hit only
    // stores a column that was not needed)
    std::vector<uint64_t> &cur_keys = cur_bits_;
    if (reused) cur_keys.assign(kKeyBits / 64, 0);
    const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
    const int nt = std::max(1, std::min({8, hw, E / 128}));
    std::vector<std::vector<int32_t>> hits(nt);
    std::vector<std::vector<int64_t>> miss(nt);  // executed positions without a column
    std::vector<std::vector<uint64_t>> bits(reused && nt > 1 ? nt : 0);
    auto hash_mark = [&](int t) {
        const int e0 = (int)((int64_t)E * t / nt), e1 = (int)((int64_t)E * (t + 1) / nt);
        uint64_t *bt = nullptr;
        if (reused) {
            if (nt > 1) bits[t].assign(kKeyBits / 64, 0);
            bt = nt > 1 ? bits[t].data() : cur_keys.data();
        }
        for (int e = e0; e < e1; ++e) {
            const int64_t a = pop.offsets[e], a1 = a + pop.lengths[e];
            strong_hash(tok + a, vals + a, pop.lengths[e], h + a, sz + a);
            for (int64_t q = a; q < a1; ++q) {
                if (sz[q] < ms) continue;
                const int s = slot_of_.find(h[q]);
                if (s >= 0) hits[t].push_back(s);
                if (bt) {
                    const uint64_t b = h[q] >> (64 - kKeyBitsLog2);
                    bt[b >> 6] |= 1ull << (b & 63);
                }
            }
            if (sz[a] >= ms) {
                const int root = lookup(a);
                if (root >= 0 && has_mse_[root]) {
                    out.src[e] = -2 - root;  // settled: nothing of it runs
                    continue;
                }
            }
            for (int64_t q = a; q < a1; ++q) {
                if (sz[q] < ms) continue;
                const int s = lookup(q);
                if (s >= 0)
                    q += sz[q] - 1;
                else if (s == -1)
                    miss[t].push_back(q);
            }
        }
    };
    if (nt == 1) {
        hash_mark(0);
    } else {
        std::vector<std::thread> pool;
        for (int t = 1; t < nt; ++t) pool.emplace_back(hash_mark, t);
        hash_mark(0);
        for (auto &th : pool) th.join();
    }
    for (int t = 0; t < nt; ++t) {
        for (int32_t s : hits[t]) seen_gen_[s] = gen_;
        if (reused && nt > 1)
            for (size_t w = 0; w < cur_keys.size(); ++w) cur_keys[w] |= bits[t][w];
    }
    bool evicted = false;
    for (int s = 0; s < used_; ++s) {
        if (live_[s] && seen_gen_[s] != gen_) {
            live_[s] = 0;
            has_mse_[s] = 0;
            free_.push_back(s);
            --live_count_;
            evicted = true;
        }
    }
    if (evicted) {
        slot_of_.reset(live_count_ + 1024);
        for (int s = 0; s < used_; ++s)
            if (live_[s]) slot_of_.put(key_of_[s], s);
    }

    ++epoch_;
    auto ready = [&](int s) { return born_epoch_[s] < epoch_; };
    auto full = [&]() { return free_.empty() && used_ >= cap_; };
    // Program for the subtree at p (n tokens): readable columns as OP_REF,
    // column-less subtrees inline, wrapped in OP_STORE when store(q) allows,
    // a slot is free and the program stays within MAX_EXPR_LEN.  root: the
    // lookup result for p.  Returns the slot that will hold p's value or -1.
    auto emit = [&](int64_t p, int n, int root, auto &&store) -> int {
        const int64_t base = (int64_t)out.tok.size();
        out.off.push_back((int32_t)base);
        int root_slot = -1;
        for (int64_t q = p; q < p + n; ++q) {
            if (sz[q] >= ms) {
                const int s = q == p ? root : lookup(q);
                if (s >= 0 && ready(s)) {
                    out.tok.push_back(OP_REF);
                    out.val.push_back((double)s);
                    if (q == p) root_slot = s;
                    q += sz[q] - 1;
                    continue;
                }
                const int64_t room = MAX_EXPR_LEN - ((int64_t)out.tok.size() - base) - (p + n - q);
                if (s == -1 && room >= 1 && store(q)) {
                    const int ns = alloc(h[q], tok + q, vals + q, sz[q]);
                    if (ns >= 0) {
                        out.tok.push_back(OP_STORE);
                        out.val.push_back((double)ns);
                        ++out.stores;
                        if (q == p) root_slot = ns;
                    }
                }
            }
            out.tok.push_back(tok[q]);
            out.val.push_back(vals[q]);
        }
        out.len.push_back((int32_t)((int64_t)out.tok.size() - base));
        out.root_slot.push_back(root_slot);
        return root_slot;
    };

    FlatMap &root_prog = root_prog_;  // root slot -> program index
    root_prog.reset(1024);

    // Store policy for a column-less subtree met once: every one, or (REUSED)
    // one whose key the previous population held.
    auto store_once = [&](int64_t q) {
        if (!reused) return true;
        if (prev_bits_.empty()) return false;
        const uint64_t b = h[q] >> (64 - kKeyBitsLog2);
        return ((prev_bits_[b >> 6] >> (b & 63)) & 1) != 0;
    };
    // Launch 1: column-less subtrees met at least twice among the executed
    // positions, largest first; a program also stores the new subtrees nested
    // in it, which are then skipped.
    {
        FlatMap &idx = occ_idx_;
        idx.reset((size_t)T / 4 + 64);
        std::vector<std::pair<int, int64_t>> occ;  // (count, first position)
        for (int t = 0; t < nt; ++t)  // expression order
            for (int64_t q : miss[t]) {
                int k = idx.find(h[q]);
                if (k < 0) {
                    k = (int)occ.size();
                    idx.put(h[q], k);
                    occ.emplace_back(0, q);
                }
                occ[k].first++;
            }
        std::vector<std::pair<int, int64_t>> cand;  // (-size, first position)
        for (auto &o : occ)
            if (o.first >= 2 && sz[o.second] < MAX_EXPR_LEN) cand.emplace_back(-sz[o.second], o.second);
        std::sort(cand.begin(), cand.end());
        for (auto &cd : cand) {
            if (full()) break;
            const int64_t q = cd.second;
            if (lookup(q) != -1) continue;  // stored by a larger program already
            // the program's root is stored (that is its purpose); nested
            // new subtrees follow the policy
            const int rs = emit(q, sz[q], -1, [&](int64_t r) { return r == q || store_once(r); });
            if (rs >= 0) root_prog.put((uint64_t)rs, (int32_t)out.off.size() - 1);
        }
    }
    if (out.off.size() > 0) {
        out.launch_begin.push_back((int32_t)out.off.size());
        ++epoch_;  // launch 1's columns are readable in launch 2
        if (launch_ready) launch_ready(out.num_launches() - 1);
    }

    // Launch 2: the expressions.
    const int first2 = (int)out.off.size();
    for (int e = 0; e < E; ++e) {
        if (out.src[e] <= -2) continue;
        const int64_t p = pop.offsets[e];
        int root = -1;
        if (sz[p] >= ms) {
            root = lookup(p);
            if (root >= 0) {
                const int pi = root_prog.find((uint64_t)root);
                if (pi >= 0) {
                    out.src[e] = pi;
                    continue;
                }
            }
        }
        out.src[e] = (int32_t)out.off.size();
        // SHARED: an expression stores nothing (a subtree met once in the
        // population is rarely met in the next one)
        const int rs = emit(p, pop.lengths[e], root,
                            [&](int64_t r) { return policy_ != SR_CACHE_GENERATIONAL_SHARED && store_once(r); });
        if (rs >= 0) root_prog.put((uint64_t)rs, out.src[e]);
    }
    if (reused) std::swap(prev_bits_, cur_bits_);
    if ((int)out.off.size() > first2) {
        out.launch_begin.push_back((int32_t)out.off.size());
        if (launch_ready) launch_ready(out.num_launches() - 1);
    }
}

void MemoPlanner::commit(const MemoPlan &p, const float *mse, const int64_t *cnt)
{
    for (int i = 0; i < p.num_programs(); ++i) {
        const int s = p.root_slot[i];
        if (s < 0) continue;
        mse_of_[s] = mse[i];
        cnt_of_[s] = cnt[i];
        has_mse_[s] = 1;
    }
}

void MemoPlanner::remembered(int s, float *mse, int64_t *cnt) const
{
    *mse = mse_of_[s];
    *cnt = cnt_of_[s];
}

// ---- evaluation ---------------------------------------------------------------
static int memo_capacity(Ctx &c, int requested)
{
    if (requested > 0) return requested;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = (size_t)1 << 34;
    const double col = (double)c.N * sizeof(float);
    const int64_t cap = (int64_t)(0.5 * (double)free_b / col);
    return (int)std::max<int64_t>(64, std::min<int64_t>(cap, 1 << 17));
}

int eval_memo(Ctx &c, const sr_population &pop, sr_stats *st, float *mse_out, int64_t *valid_out)
{
    MemoState &m = c.memo;
    hipError_t he;
    if (m.data_version != c.data_version || m.N != c.N || !m.cols.p) {
        m.cols.release();
        const int cap = memo_capacity(c, m.capacity);
        const size_t bytes = (size_t)cap * (size_t)c.N * sizeof(float);
        if ((he = hipMalloc(&m.cols.p, bytes)) != hipSuccess) {
            m.cols.p = nullptr;
            return c.hip_fail(he, "hipMalloc generational cache columns");
        }
        m.cols.bytes = bytes;
        m.planner.configure(cap, m.min_size, m.policy);
        m.data_version = c.data_version;
        m.N = c.N;
    }
    auto t0 = std::chrono::steady_clock::now();
    int rc = validate_population(c, pop);
    if (rc) return rc;
    MemoPlan &plan = m.plan;
    // Each launch is built and enqueued as soon as the planner has emitted
    // it, so launch 1 runs on the device while launch 2 is planned.  The two
    // launches use separate program buffers and result staging; the partial
    // buffers are sized for both before the first one is in flight.
    int launched = 0, res_b[2] = {0, 0}, res_n[2] = {0, 0};
    float *const cols = m.cols.as<float>();
    auto launch = [&](int k) {
        if (rc || launched >= 2) {
            if (!rc) rc = c.fail(SR_ERR_INVALID_ARG, "more than two cache launches");
            return;
        }
        const int set = launched;
        const int b = plan.launch_begin[k], n = plan.launch_begin[k + 1] - b;
        sr_population bp{n, plan.tok.data(), plan.val.data(), nullptr, plan.off.data() + b, plan.len.data() + b};
        Program pr;
        rc = build_program(c, bp, false, pr, m.planner.capacity());
        if (rc == SR_OK && set == 0) {
            LaunchPlan lp;
            rc = plan_batched(c, MODE_MSE_REF, std::max(n, pop.num_exprs), lp);
        }
        if (rc == SR_OK) rc = run_batched(c, pr, MODE_MSE_REF, cols, cols, nullptr, st, set, false);
        if (rc) return;
        HostBuf &hr = c.h_res[set];
        const size_t mbytes = ((size_t)n * 4 + 7) / 8 * 8;
        hipError_t e2 = hr.ensure(mbytes + (size_t)n * 8);
        if (e2 == hipSuccess)
            e2 = hipMemcpyAsync(hr.p, c.d_mse.p, (size_t)n * 4, hipMemcpyDeviceToHost, c.stream);
        if (e2 == hipSuccess)
            e2 = hipMemcpyAsync(hr.as<char>() + mbytes, c.d_cnt.p, (size_t)n * 8, hipMemcpyDeviceToHost,
                                c.stream);
        if (e2 == hipSuccess) e2 = hipEventRecord((set ? c.ev2 : c.ev)[3], c.stream);
        if (e2 != hipSuccess) {
            rc = c.hip_fail(e2, "generational cache evaluation");
            return;
        }
        res_b[set] = b;
        res_n[set] = n;
        ++launched;
    };
    m.planner.plan(pop, plan, launch);
    const double host_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (rc == SR_OK && launched > 0 && (he = hipStreamSynchronize(c.stream)) != hipSuccess)
        rc = c.hip_fail(he, "generational cache evaluation");
    if (rc) {
        if (launched > 0) (void)hipStreamSynchronize(c.stream);
        m.planner.forget();  // columns of this plan's slots may be unwritten
        return rc;
    }
    const int P = plan.num_programs();
    m.mse.resize(P);
    m.cnt.resize(P);
    float kernel_ms = 0.f, xfer_ms = 0.f, t = 0.f;
    for (int set = 0; set < launched; ++set) {
        const HostBuf &hr = c.h_res[set];
        const size_t mbytes = ((size_t)res_n[set] * 4 + 7) / 8 * 8;
        std::memcpy(m.mse.data() + res_b[set], hr.p, (size_t)res_n[set] * 4);
        std::memcpy(m.cnt.data() + res_b[set], hr.as<char>() + mbytes, (size_t)res_n[set] * 8);
        hipEvent_t *ev = set ? c.ev2 : c.ev;
        if (st) add_batched_timing(c, set, st);
        hipEventElapsedTime(&t, ev[0], ev[1]);
        xfer_ms += t;
        hipEventElapsedTime(&t, ev[1], ev[2]);
        kernel_ms += t;
        hipEventElapsedTime(&t, ev[2], ev[3]);
        xfer_ms += t;
    }
    m.planner.commit(plan, m.mse.data(), m.cnt.data());
    for (int e = 0; e < pop.num_exprs; ++e) {
        float v;
        int64_t k;
        const int s = plan.src[e];
        if (s >= 0) {
            v = m.mse[s];
            k = m.cnt[s];
        } else {
            m.planner.remembered(-2 - s, &v, &k);
        }
        mse_out[e] = v;
        if (valid_out) valid_out[e] = k;
    }
    if (st) {
        st->detect_time_ms += host_ms;
        st->transfer_time_ms += xfer_ms;
        st->kernel_time_ms += kernel_ms;
        st->num_subtrees = m.planner.live_slots();
        st->num_new_subtrees = (int32_t)plan.stores;
        st->num_cached = m.planner.live_slots();
        int64_t T = 0;
        for (int e = 0; e < pop.num_exprs; ++e) T += pop.lengths[e];
        st->coverage = T > 0 ? 1.0 - (double)plan.tok.size() / (double)T : 0.0;
    }
    return SR_OK;
}

}  // namespace sr

using namespace sr;

struct sr_memo_planner {
    MemoPlanner p;
    MemoPlan plan;
};

extern "C" {

int sr_cache_policy(sr_ctx *x, int32_t policy, int32_t capacity, int32_t min_size)
{
    if (!x) return SR_ERR_INVALID_ARG;
    Ctx &c = x->c;
    std::lock_guard<std::recursive_mutex> lk(c.mu);
    if (policy != SR_CACHE_REFERENCE && policy != SR_CACHE_GENERATIONAL &&
        policy != SR_CACHE_GENERATIONAL_REUSED && policy != SR_CACHE_GENERATIONAL_SHARED &&
        policy != SR_CACHE_DEVICE)
        return c.fail(SR_ERR_INVALID_ARG, "unknown cache policy");
    if (capacity < 0 || min_size < 2) return c.fail(SR_ERR_INVALID_ARG, "cache capacity < 0 or min_size < 2");
    MemoState &m = c.memo;
    m.device = policy == SR_CACHE_DEVICE;
    m.enabled = policy != SR_CACHE_REFERENCE && !m.device;
    m.capacity = capacity;
    m.min_size = min_size;
    m.policy = m.enabled ? policy : SR_CACHE_GENERATIONAL;
    m.planner.forget();
    m.cols.release();  // re-sized and re-configured at the next evaluation
    c.dm.ready = false;
    c.dm.planned = false;
    c.ga.release_graphs();
    return SR_OK;
}

int sr_memo_planner_create(int32_t capacity, int32_t min_size, int32_t policy, sr_memo_planner **out)
{
    if (!out || capacity <= 0 || min_size < 2 ||
        (policy != SR_CACHE_GENERATIONAL && policy != SR_CACHE_GENERATIONAL_REUSED &&
         policy != SR_CACHE_GENERATIONAL_SHARED))
        return SR_ERR_INVALID_ARG;
    auto *p = new sr_memo_planner;
    p->p.configure(capacity, min_size, policy);
    *out = p;
    return SR_OK;
}

int sr_memo_planner_destroy(sr_memo_planner *p)
{
    delete p;
    return SR_OK;
}

int sr_memo_plan(sr_memo_planner *p, const sr_population *pop, int32_t *num_launches, int32_t *num_programs,
                 int64_t *num_tokens)
{
    if (!p || !pop || pop->num_exprs < 0 || (pop->num_exprs > 0 && (!pop->tokens || !pop->offsets ||
                                                                     !pop->lengths ||
                                                                     (!pop->values && !pop->values_f32))))
        return SR_ERR_INVALID_ARG;
    for (int e = 0; e < pop->num_exprs; ++e) {
        const int L = pop->lengths[e];
        const int32_t *t = pop->tokens + pop->offsets[e];
        if (pop->offsets[e] < 0 || L <= 0 || L > MAX_EXPR_LEN || raw_stack_depth(t, L) < 0) return SR_ERR_BAD_EXPR;
        for (int k = 0; k < L; ++k) {
            if (t[k] == OP_REF) return SR_ERR_BAD_EXPR;
            if (t[k] == TOK_VAR) {
                double v = pop->values ? pop->values[pop->offsets[e] + k] : pop->values_f32[pop->offsets[e] + k];
                if (!(v >= 0.0)) return SR_ERR_BAD_EXPR;
            }
        }
    }
    p->p.plan(*pop, p->plan);
    if (num_launches) *num_launches = p->plan.num_launches();
    if (num_programs) *num_programs = p->plan.num_programs();
    if (num_tokens) *num_tokens = (int64_t)p->plan.tok.size();
    return SR_OK;
}

int sr_memo_plan_copy(const sr_memo_planner *p, int32_t *tokens, double *values, int32_t *offsets,
                      int32_t *lengths, int32_t *launch_begin, int32_t *root_slot, int32_t *expr_src)
{
    if (!p) return SR_ERR_INVALID_ARG;
    const MemoPlan &q = p->plan;
    if (tokens) std::memcpy(tokens, q.tok.data(), q.tok.size() * 4);
    if (values) std::memcpy(values, q.val.data(), q.val.size() * 8);
    if (offsets) std::memcpy(offsets, q.off.data(), q.off.size() * 4);
    if (lengths) std::memcpy(lengths, q.len.data(), q.len.size() * 4);
    if (launch_begin) std::memcpy(launch_begin, q.launch_begin.data(), q.launch_begin.size() * 4);
    if (root_slot) std::memcpy(root_slot, q.root_slot.data(), q.root_slot.size() * 4);
    if (expr_src) std::memcpy(expr_src, q.src.data(), q.src.size() * 4);
    return SR_OK;
}

int sr_memo_commit(sr_memo_planner *p, const float *mse, const int64_t *cnt)
{
    if (!p || (p->plan.num_programs() > 0 && (!mse || !cnt))) return SR_ERR_INVALID_ARG;
    p->p.commit(p->plan, mse, cnt);
    return SR_OK;
}

int sr_memo_remembered(const sr_memo_planner *p, int32_t slot, float *mse, int64_t *cnt)
{
    if (!p || !mse || !cnt || slot < 0 || slot >= p->p.used()) return SR_ERR_INVALID_ARG;
    p->p.remembered(slot, mse, cnt);
    return SR_OK;
}

}  // extern "C"
