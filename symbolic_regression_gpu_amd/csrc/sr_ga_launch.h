// sr_ga_launch.h -- launchers of the GA / device program-builder kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "sr_format.h"
#include "sr_ga.h"

namespace sr {

hipError_t launch_ga_init(const GaParams &p, int pop, int cap, uint64_t seed, int32_t *tok,
                          float *val, int32_t *len, int32_t *status, hipStream_t s);
// Optional history row written by the ranking launch (hist null: none).
struct GaRecord {
    uint32_t *hist = nullptr, *hist_len = nullptr;  // [max_gens][n] fitness bits, lengths
    const int32_t *len = nullptr;
    const uint32_t *gen = nullptr;
    uint32_t gen0 = 0;
    int max_gens = 0;
};
hipError_t launch_ga_rank(const float *fit, int n, int32_t *order, hipStream_t s, const GaRecord &rec = GaRecord{});
hipError_t launch_ga_next(const GaParams &p, const SelectParams &sp, int pop, int cap,
                          const int32_t *tok, const float *val, const int32_t *len, const float *fit,
                          const int32_t *order, uint64_t seed, const uint32_t *gen, int32_t *otok,
                          float *oval, int32_t *olen, int32_t *status, hipStream_t s);
hipError_t launch_compile_pop(const int32_t *tok, const float *val, const int32_t *len, int pop,
                              int cap, int nv, Ins *progs, int32_t *prog_len, int32_t *size_scratch,
                              int32_t *need_scratch, CompileFrame *frames, int32_t *max_need,
                              int32_t *status, hipStream_t s);
hipError_t launch_order_by_len(const int32_t *plen, int n, int32_t *order, hipStream_t s);
// Both orders in one launch; also zeroes zero[0, n_zero) and sets *nonzero1
// to the number of nonzero plen1 entries.
hipError_t launch_order_by_len2(const int32_t *plen0, int n0, int32_t *order0, const int32_t *plen1, int n1,
                                int32_t *order1, uint32_t *zero, int n_zero, int32_t *nonzero1, hipStream_t s);
hipError_t launch_u32_add(uint32_t *p, uint32_t v, hipStream_t s);
hipError_t launch_u32_set(uint32_t *p, uint32_t v, hipStream_t s);
hipError_t launch_evogp_to_prefix(int pop, int max_len, const float *value, const int16_t *type,
                                  const int16_t *size, int cap, int32_t *tok, float *val, int32_t *len,
                                  hipStream_t s);
hipError_t launch_prefix_to_evogp(int pop, int cap, const int32_t *tok, const float *val, const int32_t *len,
                                  int max_len, float *value, int16_t *type, int16_t *size, hipStream_t s);

}  // namespace sr

namespace sr {
hipError_t launch_splice(int n, int cap, const int32_t *ltok, const float *lval, const int32_t *llen,
                         const int32_t *lsel, const int32_t *lidx, const int32_t *rtok,
                         const float *rval, const int32_t *rlen, const int32_t *rsel,
                         const int32_t *ridx, int32_t *otok, float *oval, int32_t *olen, hipStream_t s);
}  // namespace sr
