// Batched star-stencil kernels for double grid functions.
#include "batch.h"

namespace gpaw {

hipError_t batch_launch_f64(const BatchGroup* d_groups, int ngroups, int ntiles, int R, int TZ,
                          int nt, hipStream_t st)
{
    return batch_dispatch<double>(d_groups, ngroups, ntiles, R, TZ, nt, st);
}

}  // namespace gpaw
