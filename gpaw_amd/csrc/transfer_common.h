// Point-wise restriction / interpolation shared by the transfer kernels and
// the fused coarse-level V-cycle kernel.
//
// Semantics and arithmetic follow c/bmgs/restrict.c:18-63,137-154 and
// c/bmgs/interpolate.c:19-66,141-168: three separable 1-D passes, z (axis 2)
// first, then y, then x, every pass rounded to the working precision, with
// the reference's expression grouping:
//   restrict     b = 0.5 * (a0 + w1 (a1 + a-1) + w2 (a3 + a-3) + ...)
//   interpolate  even b = a0 ; odd b = w1 (a0 + a1) + w2 (a-1 + a2) + ...
// Halo values come from ghost indexing of the input (periodic + Bloch phase
// or zero); the transformer's padding only fixes how output indices map to
// input indices:
//   restrict     output i <- input centre K-1 + 2i - pad_lo
//   interpolate  output o <- o' = o + skip_lo, i = o'/2, input K/2-1 + i - pad_lo
// Evaluating the separable passes point-wise (the z-pass values one output
// needs, then y, then x) reproduces the reference's three-sweep rounding
// exactly.
#pragma once
#include "kernels.h"

namespace gpaw {

template <int K> struct W;
template <> struct W<2> { __device__ static double w(int) { return 0.5; } };
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

// restriction taps of one axis: s[0] centre, s[2j-1] = +(2j-1), s[2j] = -(2j-1)
template <class V, int K>
__device__ __forceinline__ V rst_combine(const V* s)
{
    typedef typename VT<V>::R RT;
    V x = vmac(s[0], vadd(s[1], s[2]), (RT)W<K>::w(1));
#pragma unroll
    for (int j = 2; j <= K / 2; j++) x = vmac(x, vadd(s[2 * j - 1], s[2 * j]), (RT)W<K>::w(j));
    return vmul(x, (RT)0.5);
}

template <int K>
__device__ __forceinline__ int rst_tap(int m)   // m = 0 .. K
{
    if (m == 0) return 0;
    const int t = 2 * ((m + 1) / 2) - 1;
    return (m & 1) ? t : -t;
}

// interpolation taps of an odd output: s[0] = a0, s[1] = a1, s[2] = a-1,
// s[3] = a2, s[4] = a-2, ...
template <class V, int K>
__device__ __forceinline__ V ip_combine(const V* s)
{
    typedef typename VT<V>::R RT;
    V x = vmul(vadd(s[0], s[1]), (RT)W<K>::w(1));
#pragma unroll
    for (int j = 2; j <= K / 2; j++) x = vmac(x, vadd(s[2 * j - 2], s[2 * j - 1]), (RT)W<K>::w(j));
    return x;
}

__device__ __forceinline__ int ip_tap(int m)   // 0, 1, -1, 2, -2, ...
{
    if (m == 0) return 0;
    return (m & 1) ? (m + 1) / 2 : -(m / 2);
}

// restriction: z-pass at input row (j0, j1); the K+1 taps are unrolled
template <class V, int K>
__device__ __forceinline__ V rst_z(const V* __restrict__ a, int j0, int j1, int c2,
                                   const TrArgs& t, const Phases ph)
{
    V s[K + 1];
#pragma unroll
    for (int m = 0; m <= K; m++) s[m] = ghost_fetch(a, j0, j1, c2 + rst_tap<K>(m), t.gin, ph);
    return rst_combine<V, K>(s);
}

// y-pass over z-pass values: 0.5 * (c + w1 (+1 + -1) + w2 (+3 + -3) + ...),
// accumulated pair by pair in the reference's order
template <class V, int K>
__device__ __forceinline__ V rst_yz(const V* __restrict__ a, int j0, int c1, int c2,
                                    const TrArgs& t, const Phases ph)
{
    typedef typename VT<V>::R RT;
    V x = rst_z<V, K>(a, j0, c1, c2, t, ph);
#pragma unroll 1
    for (int j = 1; j <= K / 2; j++) {
        const int o = 2 * j - 1;
        const V p = rst_z<V, K>(a, j0, c1 + o, c2, t, ph);
        const V m = rst_z<V, K>(a, j0, c1 - o, c2, t, ph);
        x = vmac(x, vadd(p, m), (RT)W<K>::w(j));
    }
    return vmul(x, (RT)0.5);
}

template <class V, int K>
__device__ __forceinline__ V restrict_point(const V* __restrict__ a, int o0, int o1, int o2,
                                            const TrArgs& t, const Phases ph)
{
    typedef typename VT<V>::R RT;
    const int c0 = K - 1 + 2 * o0 - t.lo[0];
    const int c1 = K - 1 + 2 * o1 - t.lo[1];
    const int c2 = K - 1 + 2 * o2 - t.lo[2];
    V x = rst_yz<V, K>(a, c0, c1, c2, t, ph);
#pragma unroll 1
    for (int j = 1; j <= K / 2; j++) {
        const int o = 2 * j - 1;
        const V p = rst_yz<V, K>(a, c0 + o, c1, c2, t, ph);
        const V m = rst_yz<V, K>(a, c0 - o, c1, c2, t, ph);
        x = vmac(x, vadd(p, m), (RT)W<K>::w(j));
    }
    return vmul(x, (RT)0.5);
}

// interpolation: the odd combination w1 (a0 + a1) + w2 (a-1 + a2) + ...
template <class V, int K>
__device__ __forceinline__ V interp_z(const V* __restrict__ a, int j0, int j1, int c2, bool odd,
                                      const TrArgs& t, const Phases ph)
{
    if (!odd) return ghost_fetch(a, j0, j1, c2, t.gin, ph);
    V s[K];
#pragma unroll
    for (int m = 0; m < K; m++) s[m] = ghost_fetch(a, j0, j1, c2 + ip_tap(m), t.gin, ph);
    return ip_combine<V, K>(s);
}

template <class V, int K>
__device__ __forceinline__ V interp_yz(const V* __restrict__ a, int j0, int c1, int c2, bool oy,
                                       bool oz, const TrArgs& t, const Phases ph)
{
    typedef typename VT<V>::R RT;
    if (!oy) return interp_z<V, K>(a, j0, c1, c2, oz, t, ph);
    V x = vmul(vadd(interp_z<V, K>(a, j0, c1, c2, oz, t, ph),
                    interp_z<V, K>(a, j0, c1 + 1, c2, oz, t, ph)),
               (RT)W<K>::w(1));
#pragma unroll 1
    for (int j = 2; j <= K / 2; j++)
        x = vadd(x, vmul(vadd(interp_z<V, K>(a, j0, c1 - j + 1, c2, oz, t, ph),
                              interp_z<V, K>(a, j0, c1 + j, c2, oz, t, ph)),
                         (RT)W<K>::w(j)));
    return x;
}

template <class V, int K>
__device__ __forceinline__ V interp_point(const V* __restrict__ a, int o0, int o1, int o2,
                                          const TrArgs& t, const Phases ph)
{
    typedef typename VT<V>::R RT;
    const int p0 = o0 + t.skip[0][0], p1 = o1 + t.skip[1][0], p2 = o2 + t.skip[2][0];
    const int c0 = K / 2 - 1 + (p0 >> 1) - t.lo[0];
    const int c1 = K / 2 - 1 + (p1 >> 1) - t.lo[1];
    const int c2 = K / 2 - 1 + (p2 >> 1) - t.lo[2];
    const bool ox = p0 & 1, oy = p1 & 1, oz = p2 & 1;
    if (!ox) return interp_yz<V, K>(a, c0, c1, c2, oy, oz, t, ph);
    V x = vmul(vadd(interp_yz<V, K>(a, c0, c1, c2, oy, oz, t, ph),
                    interp_yz<V, K>(a, c0 + 1, c1, c2, oy, oz, t, ph)),
               (RT)W<K>::w(1));
#pragma unroll 1
    for (int j = 2; j <= K / 2; j++)
        x = vadd(x, vmul(vadd(interp_yz<V, K>(a, c0 - j + 1, c1, c2, oy, oz, t, ph),
                              interp_yz<V, K>(a, c0 + j, c1, c2, oy, oz, t, ph)),
                         (RT)W<K>::w(j)));
    return x;
}

// K = 2 interpolation of one coarse cell (k0, k1, k2): its 2 x 2 x 2 corner
// values produce the 2 x 2 x 2 fine points p = 2i + {0, 1} (output
// o = p - skip_lo).  Even fine coordinates copy, odd ones average
// (0.5 * (a_i + a_{i+1}), c/bmgs/interpolate.c:47-48), pass order z, y, x
// as in interp_point, so the values are bit-identical with 1/8 of the loads.
// MODE: TR_STORE (o = value) or TR_SUBTRACT (o -= value).
template <class V, int MODE>
__device__ __forceinline__ void interp2_cell(const V* __restrict__ a, V* __restrict__ o,
                                             const TrArgs& t, int k0, int k1, int k2,
                                             const Phases& ph)
{
    typedef typename VT<V>::R RT;
    // first coarse cell index per axis
    const int i0 = k0 + (t.skip[0][0] >> 1), i1 = k1 + (t.skip[1][0] >> 1),
              i2 = k2 + (t.skip[2][0] >> 1);
    const int c0 = i0 - t.lo[0], c1 = i1 - t.lo[1], c2 = i2 - t.lo[2];
    V v[2][2][2];
#pragma unroll
    for (int d0 = 0; d0 < 2; d0++)
#pragma unroll
        for (int d1 = 0; d1 < 2; d1++)
#pragma unroll
            for (int d2 = 0; d2 < 2; d2++)
                v[d0][d1][d2] = ghost_fetch(a, c0 + d0, c1 + d1, c2 + d2, t.gin, ph);
    const RT h = (RT)0.5;
    // z pass: z[d0][d1][pz]
    V z[2][2][2];
#pragma unroll
    for (int d0 = 0; d0 < 2; d0++)
#pragma unroll
        for (int d1 = 0; d1 < 2; d1++) {
            z[d0][d1][0] = v[d0][d1][0];
            z[d0][d1][1] = vmul(vadd(v[d0][d1][0], v[d0][d1][1]), h);
        }
    // y pass: y[d0][py][pz]
    V y[2][2][2];
#pragma unroll
    for (int d0 = 0; d0 < 2; d0++)
#pragma unroll
        for (int pz = 0; pz < 2; pz++) {
            y[d0][0][pz] = z[d0][0][pz];
            y[d0][1][pz] = vmul(vadd(z[d0][0][pz], z[d0][1][pz]), h);
        }
#pragma unroll
    for (int px = 0; px < 2; px++) {
        const int f0 = 2 * i0 + px - t.skip[0][0];
        if (f0 < 0 || f0 >= t.nout[0]) continue;
#pragma unroll
        for (int py = 0; py < 2; py++) {
            const int f1 = 2 * i1 + py - t.skip[1][0];
            if (f1 < 0 || f1 >= t.nout[1]) continue;
#pragma unroll
            for (int pz = 0; pz < 2; pz++) {
                const int f2 = 2 * i2 + pz - t.skip[2][0];
                if (f2 < 0 || f2 >= t.nout[2]) continue;
                const V r = px ? vmul(vadd(y[0][py][pz], y[1][py][pz]), h) : y[0][py][pz];
                const i64 q = ((i64)f0 * t.nout[1] + f1) * t.nout[2] + f2;
                if (MODE == TR_SUBTRACT)
                    o[q] = vsub(o[q], r);
                else
                    o[q] = r;
            }
        }
    }
}

// The K = 2 interpolation at a fine point of parities (px, py, pz) from its
// coarse corner values f(d0, d1, d2), d in {0, 1}: the value interp2_cell
// gives that point (z, y, x passes, same rounding).
template <class V, class F>
__device__ __forceinline__ V interp2_combine(F&& f, bool px, bool py, bool pz)
{
    typedef typename VT<V>::R RT;
    const RT h = (RT)0.5;
    V y[2];
#pragma unroll
    for (int d0 = 0; d0 < 2; d0++) {
        if (d0 && !px) break;
        V z[2];
#pragma unroll
        for (int d1 = 0; d1 < 2; d1++) {
            if (d1 && !py) break;
            const V v0 = f(d0, d1, 0);
            z[d1] = pz ? vmul(vadd(v0, f(d0, d1, 1)), h) : v0;
        }
        y[d0] = py ? vmul(vadd(z[0], z[1]), h) : z[0];
    }
    return px ? vmul(vadd(y[0], y[1]), h) : y[0];
}

// coarse cells covering every output of a K = 2 interpolation, per axis:
// i from skip_lo / 2 to (nout - 1 + skip_lo) / 2
__host__ __device__ inline int interp2_cells(const TrArgs& t, int a)
{
    return ((t.nout[a] - 1 + t.skip[a][0]) >> 1) - (t.skip[a][0] >> 1) + 1;
}

}  // namespace gpaw
