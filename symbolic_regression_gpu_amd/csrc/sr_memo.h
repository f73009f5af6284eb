// sr_memo.h -- host-side planner of the generational subtree cache.
//
// The planner owns the bookkeeping of a set of device value columns (one per
// cache slot; the columns themselves live with the caller) and turns a
// population into launches of skeleton programs for the batched interpreter:
//   * a subtree (size >= min_size) whose column was written by an earlier
//     launch is one OP_REF token;
//   * a subtree without a column is evaluated inline and, when the store
//     policy picks it and a slot is free, wrapped in OP_STORE so the launch
//     writes its column;
//   * an expression whose whole tree has a column and a remembered MSE is not
//     launched; repeats of one expression in a population are launched once.
// Launch 1 (only when needed) evaluates the column-less subtrees that occur
// more than once among the positions the expressions will execute, so that
// launch 2 (the expressions) reads them.  Slots whose subtree does not occur
// in a population are evicted before it is planned.  Keys are 64-bit
// structural hashes; every hit is verified against the stored token body.
#pragma once

#include <cstdint>
#include <functional>
#include <vector>

#include "../../include/sr_hip.h"

namespace sr {

// uint64 key -> int32 value; insert / overwrite / find, cleared in bulk.
class FlatMap {
public:
    void reset(size_t expected);
    int32_t find(uint64_t k) const;  // -1 when absent
    void put(uint64_t k, int32_t v);
    size_t size() const { return n_; }

private:
    static constexpr int32_t kEmpty = INT32_MIN;
    std::vector<uint64_t> keys_;
    std::vector<int32_t> vals_;
    size_t mask_ = 0, n_ = 0;
    void grow();
};

struct MemoPlan {
    // programs of every launch, packed; launch k holds programs
    // [launch_begin[k], launch_begin[k + 1])
    std::vector<int32_t> tok, off, len;
    std::vector<double> val;
    std::vector<int32_t> launch_begin;
    std::vector<int32_t> root_slot;  // per program: slot holding its whole value, or -1
    // per expression: index of the program giving its MSE, or -2 - slot for
    // an MSE remembered in that slot
    std::vector<int32_t> src;
    int64_t stores = 0;  // OP_STORE tokens
    int num_launches() const { return (int)launch_begin.size() - 1; }
    int num_programs() const { return (int)off.size(); }
};

class MemoPlanner {
public:
    // store_policy: SR_CACHE_GENERATIONAL (store every new subtree),
    // SR_CACHE_GENERATIONAL_REUSED (store a subtree met twice: twice among a
    // population's executed positions, or in two consecutive populations) or
    // SR_CACHE_GENERATIONAL_SHARED (store a subtree met twice in the
    // population, and one nested in such a subtree that the previous
    // population held).
    void configure(int capacity, int min_size, int store_policy);
    // Drop every slot (their columns are no longer valid).
    void forget();
    // Plan one population (validated prefix streams).  Slots handed out here
    // are written by the planned launches, which the caller must run in
    // order (or call forget()).  launch_ready(k), when given, is called as
    // soon as launch k's programs are complete (launch 0 before launch 1 is
    // planned, so the caller can start it on the device meanwhile).
    void plan(const sr_population &pop, MemoPlan &out, const std::function<void(int)> &launch_ready = {});
    // MSE and finite count of every program of the last plan.
    void commit(const MemoPlan &p, const float *mse, const int64_t *cnt);
    // MSE and finite count remembered in slot s (a plan's src = -2 - s).
    void remembered(int s, float *mse, int64_t *cnt) const;
    int capacity() const { return cap_; }
    int used() const { return used_; }
    int min_size() const { return min_size_; }
    int live_slots() const { return (int)live_count_; }

private:
    int cap_ = 0, min_size_ = 4, policy_ = SR_CACHE_GENERATIONAL;
    uint32_t gen_ = 0;
    uint64_t epoch_ = 0;  // launches planned so far
    FlatMap slot_of_;     // key -> slot, live slots only
    static constexpr int kKeyBitsLog2 = 21;
    static constexpr size_t kKeyBits = (size_t)1 << kKeyBitsLog2;
    std::vector<uint64_t> prev_bits_;  // keys met in the previous population (REUSED)
    std::vector<uint64_t> key_of_, born_epoch_;
    std::vector<uint32_t> seen_gen_;
    std::vector<uint8_t> live_, has_mse_;
    std::vector<std::vector<int32_t>> body_tok_;
    std::vector<std::vector<double>> body_val_;
    std::vector<float> mse_of_;
    std::vector<int64_t> cnt_of_;
    std::vector<int32_t> free_;
    int used_ = 0;
    size_t live_count_ = 0;
    // per-call scratch kept between calls
    FlatMap root_prog_, occ_idx_;
    std::vector<uint64_t> cur_bits_;
    std::vector<uint64_t> h_;
    std::vector<int32_t> sz_;
    std::vector<double> vals_;
    int alloc(uint64_t key, const int32_t *tok, const double *val, int n);
};

}  // namespace sr
