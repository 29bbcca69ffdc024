// k_star_tile instantiations for double2 grid functions.
#include "stencil_tile_impl.h"

namespace gpaw {

hipError_t tile_launch_c128(const Stencil& s, int mode, const void* in, void* out, const void* aux,
                          const Geom& g, const Phases& ph, double w, int TZ, const TileShape& sh,
                          double* partial, bool write, hipStream_t st)
{
    return tile_dispatch<double2>(s, mode, in, out, aux, g, ph, w, TZ, sh, partial, write, st);
}

}  // namespace gpaw
