// Restriction / interpolation dispatch (gfx950); kernels in transfer_impl.h.
#include "kernels.h"
#include "transfer_common.h"

namespace gpaw {

hipError_t transfer_launch_f64(const Transfer&, int, const void*, void*, void*, const Phases&, int,
                               hipStream_t);
hipError_t transfer_launch_c128(const Transfer&, int, const void*, void*, void*, const Phases&,
                                int, hipStream_t);
hipError_t transfer_launch_f32(const Transfer&, int, const void*, void*, void*, const Phases&, int,
                               hipStream_t);
hipError_t transfer_launch_c64(const Transfer&, int, const void*, void*, void*, const Phases&, int,
                               hipStream_t);

TrArgs make_trargs(const Transfer& tr)
{
    TrArgs t;
    t.gin = tr.gin;
    for (int a = 0; a < 3; a++) {
        t.nout[a] = tr.nout[a];
        t.lo[a] = tr.lo[a];
        t.skip[a][0] = tr.skip[a][0];
        t.skip[a][1] = tr.skip[a][1];
    }
    t.ngin = (i64)tr.gin.n[0] * tr.gin.n[1] * tr.gin.n[2];
    t.ngout = (i64)tr.nout[0] * tr.nout[1] * tr.nout[2];
    return t;
}

hipError_t launch_transfer(const Transfer& t, int dtype, int mode, const void* in, void* out,
                           void* zero_out, const Phases& ph, int nin, hipStream_t st)
{
    switch (dtype) {
    case DT_F64: return transfer_launch_f64(t, mode, in, out, zero_out, ph, nin, st);
    case DT_C128: return transfer_launch_c128(t, mode, in, out, zero_out, ph, nin, st);
    case DT_F32: return transfer_launch_f32(t, mode, in, out, zero_out, ph, nin, st);
    case DT_C64: return transfer_launch_c64(t, mode, in, out, zero_out, ph, nin, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace gpaw
