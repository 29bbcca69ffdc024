// k_star_box instantiations for float grid functions.
#include "stencil_box_impl.h"

namespace gpaw {

hipError_t box_launch_f32(const Stencil& s, int mode, const void* in, void* out, const void* aux,
                         const Geom& g, const Phases& ph, double w, const TileShape& sh,
                         double* partial, bool write, hipStream_t st)
{
    return box_dispatch<float>(s, mode, in, out, aux, g, ph, w, sh, partial, write, st);
}

hipError_t box_fz_f32(const Stencil& s, const void* rho, void* out, const Geom& g, double w0,
                      double w1, const TileShape& sh, hipStream_t st)
{
    return box_fz<float>(s, rho, out, g, w0, w1, sh, st);
}

}  // namespace gpaw
