// Vector interpreter instantiations (MSE of whole expressions and of cache
// skeletons) and their launcher.
#include "sr_launch.h"
#include "sr_vec.h"

namespace sr {

static_assert(kVecSpillSlots * kVecTile * sizeof(float) == kVecSpillBytesPerWave,
              "the spill buffer the host allocates per wave must hold the kernel's spilled slots");

hipError_t launch_vec(int mode, const BatchArgs &a, int grid, hipStream_t s)
{
    if (!a.spill) return hipErrorInvalidValue;
    if (mode == MODE_MSE)
        hipLaunchKernelGGL(k_vec<MODE_MSE>, dim3(grid), dim3(kBlock), 0, s, a);
    else if (mode == MODE_MSE_REF)
        hipLaunchKernelGGL(k_vec<MODE_MSE_REF>, dim3(grid), dim3(kBlock), 0, s, a);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

int vec_occupancy(int mode)
{
    int blocks = 0;
    const void *f = mode == MODE_MSE ? (const void *)k_vec<MODE_MSE>
                                     : (mode == MODE_MSE_REF ? (const void *)k_vec<MODE_MSE_REF> : nullptr);
    if (!f || hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, f, kBlock, 0) != hipSuccess) return 0;
    return blocks;
}

}  // namespace sr
