// Batched star-stencil kernels for float2 grid functions.
#include "batch.h"

namespace gpaw {

hipError_t batch_launch_c64(const BatchGroup* d_groups, int ngroups, int ntiles, int R, int TZ,
                          int nt, hipStream_t st)
{
    return batch_dispatch<float2>(d_groups, ngroups, ntiles, R, TZ, nt, st);
}

}  // namespace gpaw
