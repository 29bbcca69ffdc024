// Restriction / interpolation instantiations for float2 grid functions.
#include "transfer_impl.h"

namespace gpaw {

hipError_t transfer_launch_c64(const Transfer& t, int mode, const void* in, void* out,
                             void* zero_out, const Phases& ph, int nin, hipStream_t st)
{
    return tr_dispatch<float2>(t, mode, in, out, zero_out, ph, nin, st);
}

}  // namespace gpaw
