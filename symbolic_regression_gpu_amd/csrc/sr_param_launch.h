// sr_param_launch.h -- launchers of the constant-fitting kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace sr {

struct ParamArgs;

hipError_t launch_param_fwd(int dpt, int xc, const ParamArgs &a, int grid, hipStream_t s);
int param_fwd_occupancy(int dpt, int xc, int prog_lds);
hipError_t launch_adjoint(int dpt, int xc, const ParamArgs &a, int grid, hipStream_t s);
int adjoint_occupancy(int dpt, int xc, int prog_lds);
hipError_t launch_finalize_grad(const double *psum, const long long *pcnt, const double *pgrad, int P,
                                int E, float *mse, float *grad, hipStream_t s);
hipError_t launch_fd_combine(const float *cvals, const float *rep_mse, const int32_t *nslots, int E,
                             float h, float *grad, hipStream_t s);

}  // namespace sr
