// Batched star-stencil kernels for double2 grid functions.
#include "batch.h"

namespace gpaw {

hipError_t batch_launch_c128(const BatchGroup* d_groups, int ngroups, int ntiles, int R, int TZ,
                          int nt, hipStream_t st)
{
    return batch_dispatch<double2>(d_groups, ngroups, ntiles, R, TZ, nt, st);
}

}  // namespace gpaw
