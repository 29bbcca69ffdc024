// Batched star-stencil kernels for float grid functions.
#include "batch.h"

namespace gpaw {

hipError_t batch_launch_f32(const BatchGroup* d_groups, int ngroups, int ntiles, int R, int TZ,
                          int nt, hipStream_t st)
{
    return batch_dispatch<float>(d_groups, ngroups, ntiles, R, TZ, nt, st);
}

}  // namespace gpaw
