// 2x restriction and interpolation between multigrid levels (gfx950).
//
// Semantics and arithmetic follow c/bmgs/restrict.c:18-63,137-154 and
// c/bmgs/interpolate.c:19-66,141-168: three separable 1-D passes, z (axis 2)
// first, then y, then x, every pass rounded to the working precision, with
// the reference's expression grouping:
//   restrict   b = 0.5 * (a0 + w1 (a1 + a-1) + w3 (a3 + a-3) + ...)
//   interpolate even b = a0 ; odd b = w1 (a0 + a1) + w2 (a-1 + a2) + ...
// The halo comes from ghost indexing of the input (periodic + Bloch phase
// or zero), the transformer's padding only fixes how output indices map to
// input indices:
//   restrict    output i  <- input centre  K-1 + 2i - pad_lo
//   interpolate output o  <- o' = o + skip_lo, i = o'/2, input K/2-1 + i - pad_lo
// One thread produces one output point of one band.  The separable passes
// are evaluated point-wise (the z-pass values a point needs, then y, then
// x), which keeps each output bit-identical to the reference's three-sweep
// order.  TR_SUBTRACT fuses the multigrid correction phi -= interpolate(e).
#include "kernels.h"

namespace gpaw {

template <int K> struct W;
template <> struct W<2> { __device__ static double w(int j) { return 0.5; } };
template <> struct W<4> {
    __device__ static double w(int j) { return j == 1 ? 0.5625 : -0.0625; }
};
template <> struct W<6> {
    __device__ static double w(int j)
    {
        return j == 1 ? 0.58593750 : j == 2 ? -0.09765625 : 0.01171875;
    }
};
template <> struct W<8> {
    __device__ static double w(int j)
    {
        return j == 1 ? 0.59814453125 : j == 2 ? -0.11962890625 : j == 3 ? 0.02392578125
                                                                          : -0.00244140625;
    }
};

struct TrArgs {
    Geom gin;
    int nout[3];
    int lo[3];
    int skip[3][2];
    i64 ngin, ngout;
};

// ------------------------------ restriction --------------------------------
// restriction tap values along one axis: centre, then pairs (+t, -t), t = 1,3,..,2K-3
template <class V, int K>
__device__ __forceinline__ V rst_combine(const V* s)
{
    typedef typename VT<V>::R RT;
    // s[0] = centre, s[2j-1] = +(2j-1), s[2j] = -(2j-1)
    V x = vadd(s[0], vmul(vadd(s[1], s[2]), (RT)W<K>::w(1)));
#pragma unroll
    for (int j = 2; j <= K / 2; j++) x = vadd(x, vmul(vadd(s[2 * j - 1], s[2 * j]), (RT)W<K>::w(j)));
    return vmul(x, (RT)0.5);
}

template <int K>
__device__ __forceinline__ int rst_tap(int m)   // m = 0 .. K: offset of tap m
{
    if (m == 0) return 0;
    const int t = 2 * ((m + 1) / 2) - 1;
    return (m & 1) ? t : -t;
}

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
    const int c0 = K - 1 + 2 * o0 - t.lo[0];
    const int c1 = K - 1 + 2 * o1 - t.lo[1];
    const int c2 = K - 1 + 2 * o2 - t.lo[2];
    V sx[K + 1];
#pragma unroll
    for (int mx = 0; mx <= K; mx++) {
        V sy[K + 1];
#pragma unroll
        for (int my = 0; my <= K; my++) {
            V sz[K + 1];
#pragma unroll
            for (int mz = 0; mz <= K; mz++)
                sz[mz] = ghost_fetch(a, c0 + rst_tap<K>(mx), c1 + rst_tap<K>(my),
                                     c2 + rst_tap<K>(mz), t.gin, ph);
            sy[my] = rst_combine<V, K>(sz);
        }
        sx[mx] = rst_combine<V, K>(sy);
    }
    const i64 o = (i64)blockIdx.y * t.ngout + idx;
    out[o] = rst_combine<V, K>(sx);
    if (zero_out) zero_out[o] = vzero<V>();
}

// ----------------------------- interpolation -------------------------------
template <class V, int K>
__device__ __forceinline__ V ip_combine(const V* s)
{
    typedef typename VT<V>::R RT;
    // s[0] = a0, s[1] = a1, s[2] = a-1, s[3] = a2, s[4] = a-2, s[5] = a3 ...
    V x = vmul(vadd(s[0], s[1]), (RT)W<K>::w(1));
#pragma unroll
    for (int j = 2; j <= K / 2; j++) x = vadd(x, vmul(vadd(s[2 * j - 2], s[2 * j - 1]), (RT)W<K>::w(j)));
    return x;
}

// tap m = 0..K-1 of an odd output: 0, 1, -1, 2, -2, 3, ...
__device__ __forceinline__ int ip_tap(int m)
{
    if (m == 0) return 0;
    return (m & 1) ? (m + 1) / 2 : -(m / 2);
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
    const int p0 = o0 + t.skip[0][0], p1 = o1 + t.skip[1][0], p2 = o2 + t.skip[2][0];
    const int c0 = K / 2 - 1 + (p0 >> 1) - t.lo[0];
    const int c1 = K / 2 - 1 + (p1 >> 1) - t.lo[1];
    const int c2 = K / 2 - 1 + (p2 >> 1) - t.lo[2];
    const int n0 = (p0 & 1) ? K : 1, n1 = (p1 & 1) ? K : 1, n2 = (p2 & 1) ? K : 1;
    V sx[K];
    for (int mx = 0; mx < n0; mx++) {
        V sy[K];
        for (int my = 0; my < n1; my++) {
            V v;
            if (n2 == 1) {
                v = ghost_fetch(a, c0 + ip_tap(mx), c1 + ip_tap(my), c2, t.gin, ph);
            } else {
                V sz[K];
#pragma unroll
                for (int mz = 0; mz < K; mz++)
                    sz[mz] = ghost_fetch(a, c0 + ip_tap(mx), c1 + ip_tap(my), c2 + ip_tap(mz),
                                         t.gin, ph);
                v = ip_combine<V, K>(sz);
            }
            sy[my] = v;
        }
        sx[mx] = (n1 == 1) ? sy[0] : ip_combine<V, K>(sy);
    }
    const V r = (n0 == 1) ? sx[0] : ip_combine<V, K>(sx);
    const i64 o = (i64)blockIdx.y * t.ngout + idx;
    if (MODE == TR_SUBTRACT)
        out[o] = vsub(out[o], r);
    else
        out[o] = r;
}

// ------------------------------------------------------------------ host --
namespace {

template <class V, int K>
hipError_t tr_launch(const Transfer& tr, int mode, const void* in, void* out, void* zero_out,
                     const Phases& ph, int nin, hipStream_t st)
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
    dim3 grid((unsigned)((t.ngout + 255) / 256), nin, 1);
    if (!tr.interpolate) {
        hipLaunchKernelGGL((k_restrict<V, K>), grid, dim3(256), 0, st, (const V*)in, (V*)out,
                           (V*)zero_out, t, ph);
    } else if (mode == TR_SUBTRACT) {
        hipLaunchKernelGGL((k_interpolate<V, K, TR_SUBTRACT>), grid, dim3(256), 0, st,
                           (const V*)in, (V*)out, t, ph);
    } else {
        hipLaunchKernelGGL((k_interpolate<V, K, TR_STORE>), grid, dim3(256), 0, st, (const V*)in,
                           (V*)out, t, ph);
    }
    return hipGetLastError();
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

}  // namespace

hipError_t launch_transfer(const Transfer& t, int dtype, int mode, const void* in, void* out,
                           void* zero_out, const Phases& ph, int nin, hipStream_t st)
{
    switch (dtype) {
    case DT_F64: return tr_dispatch<double>(t, mode, in, out, zero_out, ph, nin, st);
    case DT_C128: return tr_dispatch<double2>(t, mode, in, out, zero_out, ph, nin, st);
    case DT_F32: return tr_dispatch<float>(t, mode, in, out, zero_out, ph, nin, st);
    case DT_C64: return tr_dispatch<float2>(t, mode, in, out, zero_out, ph, nin, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace gpaw
