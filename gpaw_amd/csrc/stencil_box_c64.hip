// k_star_box instantiations for complex64 (float2) grid functions.
#include "stencil_box_impl.h"

namespace gpaw {

hipError_t box_launch_c64(const Stencil& s, int mode, const void* in, void* out, const void* aux,
                         const Geom& g, const Phases& ph, double w, const TileShape& sh,
                         double* partial, bool write, hipStream_t st)
{
    return box_dispatch<float2>(s, mode, in, out, aux, g, ph, w, sh, partial, write, st);
}

}  // namespace gpaw
