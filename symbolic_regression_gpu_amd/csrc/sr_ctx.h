// sr_ctx.h -- the evaluator context behind the C ABI (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/sr_hip.h"
#include "sr_format.h"
#include "sr_ga.h"
#include "sr_memo.h"

namespace sr {

// Grow-only device buffer (freed with its owner).
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    hipError_t ensure(size_t need)
    {
        if (need <= bytes) return hipSuccess;
        if (p) {
            hipError_t e = hipFree(p);
            if (e != hipSuccess) return e;
            p = nullptr;
            bytes = 0;
        }
        size_t want = need + need / 4 + 256;
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess) return e;
        bytes = want;
        return hipSuccess;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <typename T>
    T *as() const { return static_cast<T *>(p); }
};

// Grow-only pinned host buffer (staging for async copies).
struct HostBuf {
    void *p = nullptr;
    size_t bytes = 0;
    HostBuf() = default;
    HostBuf(const HostBuf &) = delete;
    HostBuf &operator=(const HostBuf &) = delete;
    ~HostBuf() { release(); }
    hipError_t ensure(size_t need)
    {
        if (need <= bytes) return hipSuccess;
        if (p) {
            (void)hipHostFree(p);
            p = nullptr;
            bytes = 0;
        }
        size_t want = need + need / 4 + 256;
        hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
        if (e != hipSuccess) return e;
        bytes = want;
        return hipSuccess;
    }
    void release()
    {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <typename T>
    T *as() const { return static_cast<T *>(p); }
};

// Persistent subtree cache (cl_gpu_common.h:26-143 GPUSubtreeStateContext):
// flat-hash -> slot, value columns [capacity][N] on the device.
struct SubtreeCache {
    // Slots by flat FNV hash.  That hash keeps a difference confined to the
    // high bits of a value in the high bits (x0*x0 and x2*x2 collide), so a
    // lookup also compares the slot's body: body_tok / body_val from
    // body_off[slot], body_len[slot] tokens.
    std::unordered_multimap<uint64_t, int> slot_of;
    std::vector<uint64_t> keys;  // keys[slot]
    std::vector<int32_t> body_off, body_len, body_tok;
    std::vector<double> body_val;
    int capacity = 2048;         // cl_gpu_common.h:13 MAX_CACHED_SUBTREES
    int min_size = 3;            // detection parameters of cl_gpu_eval.cu:371-373
    int min_freq = 10;
    std::vector<int32_t> last_hints;  // cache marking of the last evaluation
    int64_t N = 0;               // columns valid for this N / X upload
    uint64_t data_version = 0;
    DevBuf cols;
    void clear()
    {
        slot_of.clear();
        keys.clear();
        body_off.clear();
        body_len.clear();
        body_tok.clear();
        body_val.clear();
    }
};

// Generational subtree cache (sr_memo.cpp): planner state and the device
// value columns [capacity][N] its slots index.
struct MemoState {
    bool enabled = false;
    bool device = false;  // SR_CACHE_DEVICE: planned on the device (sr_dmemo.hip)
    int capacity = 0;  // requested columns; 0 = sized from free HBM
    int min_size = 4;
    int policy = SR_CACHE_GENERATIONAL;
    MemoPlanner planner;
    MemoPlan plan;
    std::vector<float> mse;
    std::vector<int64_t> cnt;
    int64_t N = 0;
    uint64_t data_version = 0;
    DevBuf cols;
};

// Generational subtree cache planned on the device (sr_dmemo.hip): slot
// table, per-position planning arrays of the GA population, skeleton
// programs of both launches and running statistics, all fixed-size so that
// a generation is one captured graph.
struct DMemoState {
    bool ready = false;
    bool planned = false;  // a generation was planned and launched since setup
    int C = 0;        // slots (value columns)
    int H = 0;        // lookup table entries (power of two >= 2C)
    int H2 = 0;       // per-generation count table entries (power of two >= E * cap)
    int E = 0, cap = 0, PC = 0, L1 = 0;  // population shape, program stride, launch-1 capacity
    int64_t N = 0;
    uint64_t data_version = 0;
    int min_size = 3;
    DevBuf cols;                                                      // [C][N] fp32
    DevBuf s_key, s_len, s_tok, s_val, s_src, s_seen, s_live, s_mse, s_cnt, s_has, s_free;  // per slot
    DevBuf ht_key, ht_slot;                                           // key -> slot
    DevBuf ct_key, ct_cnt, ct_rep, ct_root, ct_slot, ct_wr, ct_off, ct_prog;  // per generation
    DevBuf sz, hh, hit, sh, anc, pslot, ctx_idx, flag, rank;          // per position [E*cap]
    DevBuf p1_slot, p1_pos, p1_need, p1_ins, p1_plen, p1_order, p1_pf;
    DevBuf p2_ins, p2_plen, p2_order, p2_pf, e_src;
    DevBuf mse1, cnt1, ctr, stats, blk;
};

// Device-resident GA population (double-buffered, fixed stride cap) and the
// buffers of its device-side evaluation.
struct GaState {
    bool ready = false;
    GaParams p{};
    int pop = 0, cap = 0, cur = 0;
    DevBuf tok[2], val[2], len[2];
    DevBuf progs, prog_len, size_s, need_s, frames, order_len, order_fit, status, max_need, gen,
        hist, hist_len, fit;
    std::vector<float> last_mse;  // host copy of the last evaluation
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    uint64_t graph_key = 0;  // parameters the captured graphs were built for
    ~GaState() { release_graphs(); }
    void release_graphs()
    {
        for (auto &g : graph)
            if (g) (void)hipGraphExecDestroy(g);
        graph[0] = graph[1] = nullptr;
        graph_key = 0;
    }
};

// Parametric population for constant fitting (free constants in a device
// array [E][8]).
struct ParamState {
    bool ready = false;
    int E = 0, max_len = 0;
    int tape_rows = 1;  // most tape rows any program needs
    int fwd_dpt = 0;    // datapoints per lane of the forward / adjoint kernels
    int adj_dpt = 0;    // (0: from the context tuning; set by sr_param_autotune)
    int64_t tokens = 0;
    std::vector<int32_t> nslots;
    DevBuf progs, aux, meta, psum, pcnt, pgrad, tape, fd_mse;
};

// hipRTC module of the last JIT-evaluated population.
struct JitState {
    hipModule_t mod = nullptr;
    hipFunction_t fn = nullptr;
    // The code object the module was loaded from.  The runtime may read the
    // image after hipModuleLoadData returns (deferred loading), so it lives as
    // long as the module.
    std::vector<char> image;
    uint64_t key = 0;
    int occ = 1;
    int64_t code_bytes = 0;
    double last_compile_ms = 0.0;
    JitState() = default;
    JitState(const JitState &) = delete;
    JitState &operator=(const JitState &) = delete;
    ~JitState();
};

// Device address of the static scratch array of sr_eval_mse_noalloc on
// `device` (sr_noalloc.hip); call with that device current.
hipError_t noalloc_scratch_init(int device, void **out);

struct Ctx {
    int device = 0;
    void *noalloc_scratch = nullptr;  // this device's static scratch (sr_eval_mse_noalloc)
    hipStream_t stream = nullptr;
    hipStream_t own_stream = nullptr;
    int num_cu = 0;
    char arch[64] = {0};  // gcnArchName without target features (e.g. "gfx950")
    std::string err;
    std::recursive_mutex mu;

    // X|y
    const float *X = nullptr;
    DevBuf X_own;
    int nv = 0;
    int64_t N = 0;
    uint64_t data_version = 0;

    sr_tuning tune{3, 3072, 8, 0, 3, 1, 1};

    // staging (host pinned) and device buffers for one evaluation; the
    // second program set lets a launch be prepared while one is in flight
    HostBuf h_prog, h_meta, h_out;
    DevBuf d_prog, d_meta, d_psum, d_pcnt, d_mse, d_cnt, d_counters, d_pred, d_spill;
    HostBuf h_prog2, h_meta2, h_res[2];
    DevBuf d_prog2, d_meta2;

    SubtreeCache cache;
    MemoState memo;
    GaState ga;
    DMemoState dm;
    ParamState par;
    JitState jit;
    // ev[0..3]: transfer / kernel / transfer brackets of one call;
    // ev[4..5]: the interpreter kernel alone.
    hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    hipEvent_t ev2[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};  // program set 1

    int fail(int code, const std::string &msg)
    {
        err = msg;
        return code;
    }
    int hip_fail(hipError_t e, const char *what)
    {
        err = std::string(what) + ": " + hipGetErrorString(e);
        return SR_ERR_HIP;
    }
};

// Host-side packed program for a population.
struct Program {
    std::vector<Ins> ins;
    std::vector<int32_t> off, len, order, need;
    std::vector<int32_t> ref_pf;  // [E][kRefPrefetch] (cache skeletons only)
    int max_need = 0;
    int64_t tokens = 0;
    int64_t ops = 0, refs = 0, stores = 0;  // operator / OP_REF / OP_STORE tokens
};

// Validate and compile every expression of pop into bytecode (batched) or
// lowered raw tokens (raw = true, plain interpreter).  Returns SR_OK or
// SR_ERR_BAD_EXPR with ctx.err set.
int build_program(Ctx &c, const sr_population &pop, bool raw, Program &out, int num_slots = 0);
// Values of expression e as float (values_f32 or (float)values).
void expr_values_f32(const sr_population &pop, int e, std::vector<float> &out);

// Launch shape of one batched evaluation (see plan_batched).
struct LaunchPlan {
    int dpt = 4, xc = 4, nq = 8, cpq = 1, P = 8, grid = 0;
    int kind = 0;  // KERNEL_LANE or KERNEL_VEC
    int64_t slice = 0, chunk = 0;
};
struct BatchArgs;
// Pick the variant + work decomposition (no allocation); allow_vec = false
// keeps the lane interpreter (no spill area).
int plan_launch(Ctx &c, int mode, LaunchPlan &lp, bool allow_vec = true);
// plan_launch for E expressions, and make sure the context's partial /
// result buffers exist.
int plan_batched(Ctx &c, int mode, int E, LaunchPlan &lp);
// Device buffers one batched evaluation writes: MSE partials [E][P], the
// per-XCD work counters, and (MSE modes) the finalized mse / count [E].
struct EvalBufs {
    double *psum;
    long long *pcnt;
    uint32_t *counters;
    float *mse;
    long long *cnt;
    const int *fin_count = nullptr;  // finalize only programs below this device count
    float *spill = nullptr;          // vector interpreter stack spill (plan kind KERNEL_VEC)
    bool counters_zeroed = false;    // the caller zeroed the queue counters already
    bool finalize = true;            // MSE modes: reduce the partials (else the caller does)
};
// Enqueue counter reset + batched kernel (+ finalize for MODE_MSE).  a must
// carry the program / order / output pointers; the rest is filled in here.
// k0/k1 (optional) are recorded right before and after the interpreter launch.
int enqueue_batched(Ctx &c, int mode, const LaunchPlan &lp, BatchArgs a, int E, hipStream_t s,
                    hipEvent_t k0 = nullptr, hipEvent_t k1 = nullptr);
// Same into explicit buffers (the context's own are not touched).
int enqueue_batched(Ctx &c, int mode, const LaunchPlan &lp, BatchArgs a, int E, hipStream_t s,
                    const EvalBufs &b, hipEvent_t k0 = nullptr, hipEvent_t k1 = nullptr);

// Batched evaluation over a compiled program already described by host
// vectors; uploads, launches, finalizes.  mode: MODE_MSE/MODE_PRED/MODE_COLUMN.
// For MODE_MSE writes d_mse/d_cnt on the device (caller copies back).
// set selects the staging / program buffers and events (0 or 1); with
// wait = false the interpreter time is not read here: call
// add_batched_timing once the stream has synchronised.
int run_batched(Ctx &c, const Program &p, int mode, const float *cache_cols, float *out_dev,
                const int32_t *out_col_host, sr_stats *st, int set = 0, bool wait = true);
void add_batched_timing(Ctx &c, int set, sr_stats *st);
// Same checks as build_program without compiling (lengths, stream shape,
// variable indices; OP_REF / OP_STORE rejected).
int validate_population(Ctx &c, const sr_population &pop);

// Device-planned generational cache (sr_dmemo.hip): size it for the GA
// population and data, enqueue the evaluation of GA buffer b (fitness into
// c.d_mse), and read its statistics (see dmemo_stats).
int dmemo_setup(Ctx &c, int capacity, int min_size);
int dmemo_enqueue(Ctx &c, int b, hipStream_t s, long long *cnt_out, hipEvent_t *ev = nullptr);
int dmemo_stats(Ctx &c, long long out[9]);
// The two interpreter launches of the last planned generation again (same
// columns, same partial sums); ev: optional {start1, end1, start2, end2}.
// Within dmemo_enqueue (planned = true) the queue counters are already zero
// and launch 2's partials are reduced by the fitness gather (*P_out: the
// partials per program).
int dmemo_launch(Ctx &c, hipStream_t s, hipEvent_t *ev = nullptr, bool planned = false, int *P_out = nullptr);

}  // namespace sr

struct sr_ctx {
    sr::Ctx c;
};
