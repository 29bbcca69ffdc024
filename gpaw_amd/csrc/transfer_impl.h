// Restriction / interpolation kernels and their launchers, instantiated
// once per storage type in transfer_<type>.hip.
//
// One thread produces one output point of one band; the arithmetic is the
// point-wise separable evaluation of transfer_common.h, bit-identical to the
// reference's three-sweep bmgs_restrict / bmgs_interpolate
// (c/bmgs/restrict.c:137-154, c/bmgs/interpolate.c:141-168).
// TR_SUBTRACT fuses the multigrid correction phi -= interpolate(e)
// (gpaw/poisson.py:263-264), and restriction can zero the next level's phi
// in the same pass (gpaw/poisson.py:259-261).
#pragma once
#include <cstdlib>
#include "kernels.h"
#include "transfer_common.h"
#include "transfer_tile.h"

namespace gpaw {

TrArgs make_trargs(const Transfer& tr);

template <class V, int K>
__global__ __launch_bounds__(256) void k_restrict(const V* __restrict__ in, V* __restrict__ out,
                                                  V* __restrict__ zero_out, TrArgs t, Phases ph)
{
    const i64 idx = (i64)blockIdx.x * 256 + threadIdx.x;
    if (idx >= t.ngout) return;
    const int o2 = (int)(idx % t.nout[2]);
    const int o1 = (int)((idx / t.nout[2]) % t.nout[1]);
    const int o0 = (int)(idx / ((i64)t.nout[1] * t.nout[2]));
    const V* __restrict__ a = in + (i64)blockIdx.y * t.ngin;
    const i64 o = (i64)blockIdx.y * t.ngout + idx;
    out[o] = restrict_point<V, K>(a, o0, o1, o2, t, ph);
    if (zero_out) zero_out[o] = vzero<V>();
}

template <class V, int K, int MODE>
__global__ __launch_bounds__(256) void k_interpolate(const V* __restrict__ in, V* __restrict__ out,
                                                     TrArgs t, Phases ph)
{
    const i64 idx = (i64)blockIdx.x * 256 + threadIdx.x;
    if (idx >= t.ngout) return;
    const int o2 = (int)(idx % t.nout[2]);
    const int o1 = (int)((idx / t.nout[2]) % t.nout[1]);
    const int o0 = (int)(idx / ((i64)t.nout[1] * t.nout[2]));
    const V* __restrict__ a = in + (i64)blockIdx.y * t.ngin;
    const V r = interp_point<V, K>(a, o0, o1, o2, t, ph);
    const i64 o = (i64)blockIdx.y * t.ngout + idx;
    if (MODE == TR_SUBTRACT)
        out[o] = vsub(out[o], r);
    else
        out[o] = r;
}

// K = 2 interpolation, one thread per coarse cell (interp2_cell).
template <class V, int MODE>
__global__ __launch_bounds__(256) void k_interp2(const V* __restrict__ in, V* __restrict__ out,
                                                 TrArgs t, int nc0, int nc1, int nc2, Phases ph)
{
    const i64 idx = (i64)blockIdx.x * 256 + threadIdx.x;
    const i64 ncell = (i64)nc0 * nc1 * nc2;
    if (idx >= ncell) return;
    const int k2 = (int)(idx % nc2);
    const int k1 = (int)((idx / nc2) % nc1);
    const int k0 = (int)(idx / ((i64)nc1 * nc2));
    interp2_cell<V, MODE>(in + (i64)blockIdx.y * t.ngin, out + (i64)blockIdx.y * t.ngout, t, k0,
                          k1, k2, ph);
}

template <class V, int K>
hipError_t tr_launch(const Transfer& tr, int mode, const void* in, void* out, void* zero_out,
                     const Phases& ph, int nin, hipStream_t st)
{
    const TrArgs t = make_trargs(tr);
    dim3 grid((unsigned)((t.ngout + 255) / 256), nin, 1);
    if (K == 2 && tr.interpolate) {
        // coarse cells covering every output: i from skip_lo/2 to (nout-1+skip_lo)/2
        int nc[3];
        for (int a = 0; a < 3; a++) nc[a] = interp2_cells(t, a);
        const i64 ncell = (i64)nc[0] * nc[1] * nc[2];
        dim3 g2((unsigned)((ncell + 255) / 256), nin, 1);
        if (mode == TR_SUBTRACT)
            launch_kernel((k_interp2<V, TR_SUBTRACT>), g2, dim3(256), 0, st, (const V*)in,
                               (V*)out, t, nc[0], nc[1], nc[2], ph);
        else
            launch_kernel((k_interp2<V, TR_STORE>), g2, dim3(256), 0, st, (const V*)in,
                               (V*)out, t, nc[0], nc[1], nc[2], ph);
        return launch_error();
    }
    // wide transfers on large planes, and K = 2 restrictions of many points
    // (band sets: the point-wise kernel reads each input up to 8 times through
    // the caches): LDS-tiled plane sweep (8 x 32 tiles)
    const bool big2 = K == 2 && !tr.interpolate && t.ngout * nin >= (1 << 20);
    if ((K >= 4 || big2) && t.nout[1] >= 4 && t.nout[2] >= 16 &&
        (long)t.gin.n[1] * t.gin.n[2] < PLANE_REF_MAX && !getenv("GPAW_AMD_POINTWISE_TRANSFER")) {
        TileTr tt;
        tt.nt1 = (t.nout[1] + 7) / 8;
        tt.nt2 = (t.nout[2] + 31) / 32;
        dim3 gt((unsigned)(tt.nt1 * tt.nt2 * nin));
        if (!tr.interpolate)
            launch_kernel((k_restrict_tile<V, K, 8, 32>), gt, dim3(256), 0, st, (const V*)in,
                               (V*)out, t, tt, ph);
        else if (mode == TR_SUBTRACT)
            launch_kernel((k_interp_tile<V, K, 8, 32, TR_SUBTRACT>), gt, dim3(256), 0, st,
                               (const V*)in, (V*)out, t, tt, ph);
        else
            launch_kernel((k_interp_tile<V, K, 8, 32, TR_STORE>), gt, dim3(256), 0, st,
                               (const V*)in, (V*)out, t, tt, ph);
        if (!tr.interpolate && zero_out) return launch_fill(zero_out, VT<V>::cplx ? (sizeof(typename VT<V>::R) == 8 ? DT_C128 : DT_C64) : (sizeof(typename VT<V>::R) == 8 ? DT_F64 : DT_F32), t.ngout * nin, st);
        return launch_error();
    }
    if (!tr.interpolate) {
        launch_kernel((k_restrict<V, K>), grid, dim3(256), 0, st, (const V*)in, (V*)out,
                           (V*)zero_out, t, ph);
    } else if (mode == TR_SUBTRACT) {
        launch_kernel((k_interpolate<V, K, TR_SUBTRACT>), grid, dim3(256), 0, st,
                           (const V*)in, (V*)out, t, ph);
    } else {
        launch_kernel((k_interpolate<V, K, TR_STORE>), grid, dim3(256), 0, st, (const V*)in,
                           (V*)out, t, ph);
    }
    return launch_error();
}

template <class V>
hipError_t tr_dispatch(const Transfer& tr, int mode, const void* in, void* out, void* zero_out,
                       const Phases& ph, int nin, hipStream_t st)
{
    switch (tr.K) {
    case 2: return tr_launch<V, 2>(tr, mode, in, out, zero_out, ph, nin, st);
    case 4: return tr_launch<V, 4>(tr, mode, in, out, zero_out, ph, nin, st);
    case 6: return tr_launch<V, 6>(tr, mode, in, out, zero_out, ph, nin, st);
    case 8: return tr_launch<V, 8>(tr, mode, in, out, zero_out, ph, nin, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace gpaw
