// sr_launch.h -- host-side launchers for the evaluation kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "sr_format.h"

namespace sr {

struct BatchArgs;

constexpr int kStackSlots = 8;  // bytecode stack need <= 8 for any tree of <= 512 leaves
// Registers allotted to each datapoint's stack.  One more than the 8 slots a
// program may use, on purpose: the backend reads / writes a VGPR array of <= 8
// floats at a uniform dynamic index through a compare + v_cndmask chain over
// every element (16 VALU per access for 8 slots), and larger arrays through
// GPR-index mode (one v_mov).
constexpr int kStackRegs = 9;

// Batched persistent interpreter.  mode: 0 MSE partials, 1 predictions,
// 2 cache columns.  Returns hipErrorInvalidValue for an unsupported (dpt, xc).
hipError_t launch_batched(int mode, int dpt, int xc, const BatchArgs &a, int grid,
                          hipStream_t s);
// Max resident workgroups per CU for a variant (0 if unsupported).
int batched_occupancy(int mode, int dpt, int xc);
// Vector interpreter (sr_vec.h), MODE_MSE and MODE_MSE_REF: a.spill must
// hold kVecSpillBytesPerWave per launched wave.
hipError_t launch_vec(int mode, const BatchArgs &a, int grid, hipStream_t s);
int vec_occupancy(int mode);
// 4 spilled stack slots x 384 datapoints x 4 bytes
constexpr size_t kVecSpillBytesPerWave = (size_t)4 * 384 * 4;
// Interpreter kinds of a launch plan.
enum { KERNEL_LANE = 0, KERNEL_VEC = 1 };

// Plain interpreter, one expression.  s_slots >= raw stack depth.
hipError_t launch_plain(int s_slots, int xc, const Ins *prog, int len, const float *X,
                        int64_t N, int nv, double *psum, long long *pcnt, hipStream_t s);

// mse[e] = sum / count over P partials (NaN when count == 0), for e < E
// (and e < *E_dev when E_dev is set).
hipError_t launch_finalize(const double *psum, const long long *pcnt, int P, int E,
                           float *mse, long long *cnt, hipStream_t s, const int *E_dev = nullptr);
// Zero n words (graph-capturable: a kernel, not a memset node).
hipError_t launch_zero_u32(uint32_t *p, int n, hipStream_t s);

// Smallest register-cached variable count that holds nv variables.
inline int pick_xc(int nv) { return nv <= 4 ? 4 : (nv <= 16 ? 16 : 32); }
// The batched interpreter also has a 3-variable class for its 3-datapoint
// variant (a variable operand is then 2 selects per datapoint instead of 3).
inline int pick_xc_batched(int nv, int dpt) { return nv <= 3 && dpt == 3 ? 3 : pick_xc(nv); }

}  // namespace sr
