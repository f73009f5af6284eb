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

namespace sr {

// Grow-only device buffer.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
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
    std::unordered_map<uint64_t, int> slot_of;
    std::vector<uint64_t> keys;  // keys[slot]
    int capacity = 2048;         // cl_gpu_common.h:13 MAX_CACHED_SUBTREES
    int64_t N = 0;               // columns valid for this N / X upload
    uint64_t data_version = 0;
    DevBuf cols;
    void clear()
    {
        slot_of.clear();
        keys.clear();
    }
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t own_stream = nullptr;
    int num_cu = 0;
    std::string err;
    std::recursive_mutex mu;

    // X|y
    const float *X = nullptr;
    DevBuf X_own;
    int nv = 0;
    int64_t N = 0;
    uint64_t data_version = 0;

    sr_tuning tune{4, 16384, 8, 0};

    // staging (host pinned) and device buffers for one evaluation
    HostBuf h_prog, h_meta, h_out;
    DevBuf d_prog, d_meta, d_psum, d_pcnt, d_mse, d_cnt, d_counters, d_pred;

    SubtreeCache cache;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};

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
    int max_need = 0;
    int64_t tokens = 0;
};

// Validate and compile every expression of pop into bytecode (batched) or
// lowered raw tokens (raw = true, plain interpreter).  Returns SR_OK or
// SR_ERR_BAD_EXPR with ctx.err set.
int build_program(Ctx &c, const sr_population &pop, bool raw, Program &out, int num_slots = 0);
// Values of expression e as float (values_f32 or (float)values).
void expr_values_f32(const sr_population &pop, int e, std::vector<float> &out);

// Batched evaluation over a compiled program already described by host
// vectors; uploads, launches, finalizes.  mode: MODE_MSE/MODE_PRED/MODE_COLUMN.
// For MODE_MSE writes d_mse/d_cnt on the device (caller copies back).
int run_batched(Ctx &c, const Program &p, int mode, const float *cache_cols, float *out_dev,
                const int32_t *out_col_host, sr_stats *st);

}  // namespace sr

struct sr_ctx {
    sr::Ctx c;
};
