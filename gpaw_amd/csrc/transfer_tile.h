// LDS-tiled restriction / interpolation for large grids and wide stencils
// (gfx950), used for K >= 4 where the point-wise kernels would re-read each
// input many times.
//
// Both kernels own a tile of the output (i1, i2) plane (256 threads, one
// output column each) and sweep the input planes along i0 in order:
//   1. the tile's slab of the current input plane (with its halo) is stored
//      into LDS from registers, ghost-indexed (periodic + Bloch phase or
//      zero); the next plane's slab is loaded into those registers while
//      the current one is combined (restriction);
//   2. the z-pass is evaluated from LDS into a second LDS array;
//   3. the y-pass is evaluated from that, one value per thread, and pushed
//      into a per-thread register ring of the last K (interpolation) or
//      2K-1 (restriction) input planes;
//   4. whenever the ring holds all x-taps of an output plane, the x-pass is
//      evaluated from the ring and stored.
// Every pass uses the reference's expression grouping and pass order
// (z, y, x), so results are bit-identical to c/bmgs/restrict.c and
// c/bmgs/interpolate.c.  Slab cells outside the transformer's padded input
// region are never used by a valid output and are set to zero without a
// load, which keeps every ghost index within one period.
#pragma once
#include "transfer_common.h"

namespace gpaw {

struct TileTr {
    int nt1, nt2;   // output tiles along i1, i2
};

// ------------------------------------------------------------ restriction
template <class V, int K, int T1, int T2>
__global__ __launch_bounds__(256) void k_restrict_tile(const V* __restrict__ in, V* __restrict__ out,
                                                       TrArgs t, TileTr tt, Phases ph)
{
    constexpr int SR = 2 * T1 + 2 * K - 3;   // input rows of the slab
    constexpr int SC = 2 * T2 + 2 * K - 3;   // input columns of the slab
    constexpr int W = 2 * K - 1;             // x window
    __shared__ V A[SR * SC];
    __shared__ V Z[SR * T2];
    const int b = xcd_tile(blockIdx.x, gridDim.x);
    const int tz = b % tt.nt2;
    const int ty = (b / tt.nt2) % tt.nt1;
    const int band = b / (tt.nt2 * tt.nt1);
    const int tid = threadIdx.x;
    const int tzi = tid % T2, tyi = tid / T2;
    const V* __restrict__ a = in + (i64)band * t.ngin;
    V* __restrict__ o = out + (i64)band * t.ngout;
    const int o1 = ty * T1 + tyi, o2 = tz * T2 + tzi;
    // slab origin (input index of slab row/col 0)
    const int r0 = 2 * ty * T1 - t.lo[1];
    const int s0 = 2 * tz * T2 - t.lo[2];
    const int lo0 = -t.lo[0], lo1 = -t.lo[1], lo2 = -t.lo[2];
    const int hi0 = 2 * t.nout[0] + 2 * K - 3 - t.lo[0];
    const int hi1 = 2 * t.nout[1] + 2 * K - 3 - t.lo[1];
    const int hi2 = 2 * t.nout[2] + 2 * K - 3 - t.lo[2];
    const int p0s = -t.lo[0];
    const int np = 2 * t.nout[0] + 2 * K - 3;   // input planes swept
    // slab cells of this thread: c = tid + 256 k, with their in-plane
    // references (code 3: a cell no valid output uses, zero without a
    // load).  The next plane's cells are loaded raw into registers while the
    // current plane is combined and fixed (zero / phase) when stored to LDS.
    constexpr int NS = (SR * SC + 255) / 256;
    auto cref = [&](int k) -> PlaneRef {   // recomputed: registers are scarce here
        const int c = tid + 256 * k;
        const int j1 = r0 + c / SC, j2 = s0 + c % SC;
        const bool ok = c < SR * SC && j1 >= lo1 && j1 < hi1 && j2 >= lo2 && j2 < hi2;
        return ok ? plane_ref(j1, j2, t.gin) : PREF_ZERO;
    };
    const i64 plane = (i64)t.gin.n[1] * t.gin.n[2];
    auto pflag = [&](int p0, i64& base) -> int {
        int f = 3;
        if (p0 >= lo0 && p0 < hi0) f = wrap_index(p0, t.gin.n[0], t.gin.per[0]);
        base = f == 3 ? 0 : (i64)p0 * plane;
        return f;
    };
    V pre[NS];
    int pf;
    {
        i64 base;
        pf = pflag(p0s, base);
#pragma unroll
        for (int k = 0; k < NS; k++) pre[k] = a[base + pref_off(cref(k))];
    }
    const bool own = o1 < t.nout[1] && o2 < t.nout[2];
    V ring[W];
    // x-pass at the start of plane step j, over the ring as step j - 1 left
    // it: output plane o0 is complete when its last tap plane arrived.  Its
    // store is issued with the next plane's loads.
    auto xpass = [&](int j) {
        if (j >= W && ((j - W) & 1) == 0) {
            V s[K + 1];
#pragma unroll
            for (int m = 0; m <= K; m++) s[m] = ring[K - 1 + rst_tap<K>(m)];
            const V r = rst_combine<V, K>(s);
            if (own) o[((i64)((j - W) >> 1) * t.nout[1] + o1) * t.nout[2] + o2] = r;
        }
    };
    for (int j = 0; j < np; j++) {
        // one wait per plane step, for loads issued a step earlier
#pragma unroll
        for (int k = 0; k < NS; k++) vtouch(pre[k]);
#pragma unroll
        for (int k = 0; k < NS; k++)
            if (tid + 256 * k < SR * SC) A[tid + 256 * k] = ghost_fix(pre[k], pf, cref(k), ph);
        xpass(j);
        if (j + 1 < np) {
            i64 base;
            pf = pflag(p0s + j + 1, base);
#pragma unroll
            for (int k = 0; k < NS; k++) pre[k] = a[base + pref_off(cref(k))];
        }
        __syncthreads();
        for (int c = tid; c < SR * T2; c += 256) {
            const int r = c / T2, z = c % T2;
            V s[K + 1];
#pragma unroll
            for (int m = 0; m <= K; m++) s[m] = A[r * SC + 2 * z + K - 1 + rst_tap<K>(m)];
            Z[c] = rst_combine<V, K>(s);
        }
        __syncthreads();
        V y;
        {
            V s[K + 1];
#pragma unroll
            for (int m = 0; m <= K; m++) s[m] = Z[(2 * tyi + K - 1 + rst_tap<K>(m)) * T2 + tzi];
            y = rst_combine<V, K>(s);
        }
#pragma unroll
        for (int m = 0; m < W - 1; m++) ring[m] = ring[m + 1];
        ring[W - 1] = y;
    }
    xpass(np);
}

// ----------------------------------------------------------- interpolation
template <class V, int K, int T1, int T2, int MODE>
__global__ __launch_bounds__(256) void k_interp_tile(const V* __restrict__ in, V* __restrict__ out,
                                                     TrArgs t, TileTr tt, Phases ph)
{
    constexpr int SR = T1 / 2 + K + 1;   // coarse rows of the slab (upper bound)
    constexpr int SC = T2 / 2 + K + 1;
    __shared__ V A[SR * SC];
    __shared__ V Z[SR * T2];
    const int b = xcd_tile(blockIdx.x, gridDim.x);
    const int tz = b % tt.nt2;
    const int ty = (b / tt.nt2) % tt.nt1;
    const int band = b / (tt.nt2 * tt.nt1);
    const int tid = threadIdx.x;
    const int tzi = tid % T2, tyi = tid / T2;
    const V* __restrict__ a = in + (i64)band * t.ngin;
    V* __restrict__ o = out + (i64)band * t.ngout;
    const int o1 = ty * T1 + tyi, o2 = tz * T2 + tzi;
    // coarse centre index of fine output o along an axis
    auto centre = [&](int o, int ax) { return K / 2 - 1 + ((o + t.skip[ax][0]) >> 1) - t.lo[ax]; };
    // slab origin: coarse row / column of slab index 0
    const int r0 = centre(ty * T1, 1) - K / 2 + 1;
    const int s0 = centre(tz * T2, 2) - K / 2 + 1;
    // coarse indices a valid output can touch: the padded input region
    const int lo1 = -t.lo[1], lo2 = -t.lo[2];
    const int hi1 = centre(t.nout[1] - 1, 1) + K / 2 + 1;
    const int hi2 = centre(t.nout[2] - 1, 2) + K / 2 + 1;
    const int c1 = centre(o1, 1), c2 = centre(o2, 2);
    const bool oy = (o1 + t.skip[1][0]) & 1, oz = (o2 + t.skip[2][0]) & 1;
    // z-pass target of this thread's Z entries: fine column z of the tile
    const int c0f = centre(0, 0);                         // first output's centre
    const int p0s = c0f - K / 2 + 1;                      // first plane swept
    const int p0e = centre(t.nout[0] - 1, 0) + K / 2;     // last plane swept
    // coarse slab cells of this thread (c = tid + 256 k) with their in-plane
    // references; loaded raw one plane ahead, fixed when stored to LDS
    constexpr int NSI = (SR * SC + 255) / 256;
    auto cref = [&](int k) -> PlaneRef {
        const int c = tid + 256 * k;
        const int j1 = r0 + c / SC, j2 = s0 + c % SC;
        const bool ok = c < SR * SC && j1 >= lo1 && j1 < hi1 && j2 >= lo2 && j2 < hi2;
        return ok ? plane_ref(j1, j2, t.gin) : PREF_ZERO;
    };
    const i64 plane = (i64)t.gin.n[1] * t.gin.n[2];
    auto pflag = [&](int p0, i64& base) -> int {
        const int f = wrap_index(p0, t.gin.n[0], t.gin.per[0]);
        base = f == 3 ? 0 : (i64)p0 * plane;
        return f;
    };
    const bool own = o1 < t.nout[1] && o2 < t.nout[2];
    // index of this thread's fine output of parity par among the outputs
    // with coarse centre c0 (-1: none)
    auto outq = [&](int c0, int par) -> i64 {
        const int f0 = 2 * (c0 - (K / 2 - 1) + t.lo[0]) + par - t.skip[0][0];
        return (own && f0 >= 0 && f0 < t.nout[0]) ? ((i64)f0 * t.nout[1] + o1) * t.nout[2] + o2
                                                   : -1;
    };
    V pre[NSI];
    int pf;
    {
        i64 base;
        pf = pflag(p0s, base);
#pragma unroll
        for (int k = 0; k < NSI; k++) pre[k] = a[base + pref_off(cref(k))];
    }
    V ring[K];
    V old[2] = {vzero<V>(), vzero<V>()};   // subtract: outputs' values, loaded a step ahead
    // outputs completed by step p0 - 1 (fine planes whose coarse centre is
    // p0 - 1 - K/2, even and odd), written at the start of step p0 so that
    // their stores are issued together with the next plane's loads
    auto opass = [&](int p0) {
        if (p0 - p0s >= K) {
            const int c0 = p0 - 1 - K / 2;
#pragma unroll
            for (int par = 0; par < 2; par++) {
                const i64 q = outq(c0, par);
                if (q < 0) continue;
                V r;
                if (!par) {
                    r = ring[K / 2 - 1];
                } else {
                    V s[K];
#pragma unroll
                    for (int m = 0; m < K; m++) s[m] = ring[K / 2 - 1 + ip_tap(m)];
                    r = ip_combine<V, K>(s);
                }
                o[q] = MODE == TR_SUBTRACT ? vsub(old[par], r) : r;
            }
        }
    };
    for (int p0 = p0s; p0 <= p0e; p0++) {
        // one wait per plane step, for loads issued a step earlier
#pragma unroll
        for (int k = 0; k < NSI; k++) vtouch(pre[k]);
        if (MODE == TR_SUBTRACT) {
            vtouch(old[0]);
            vtouch(old[1]);
        }
#pragma unroll
        for (int k = 0; k < NSI; k++)
            if (tid + 256 * k < SR * SC) A[tid + 256 * k] = ghost_fix(pre[k], pf, cref(k), ph);
        opass(p0);
        if (p0 < p0e) {
            i64 base;
            pf = pflag(p0 + 1, base);
#pragma unroll
            for (int k = 0; k < NSI; k++) pre[k] = a[base + pref_off(cref(k))];
        }
        if (MODE == TR_SUBTRACT && p0 + 1 - p0s >= K) {
#pragma unroll
            for (int par = 0; par < 2; par++) {
                const i64 q = outq(p0 - K / 2, par);
                old[par] = o[q < 0 ? 0 : q];
            }
        }
        __syncthreads();
        for (int c = tid; c < SR * T2; c += 256) {
            const int r = c / T2, z = c % T2;
            const int fz = tz * T2 + z;
            const int cc = centre(fz, 2) - s0;              // slab column of the centre
            V v;
            if (!((fz + t.skip[2][0]) & 1)) {
                v = A[r * SC + cc];
            } else {
                V s[K];
#pragma unroll
                for (int m = 0; m < K; m++) s[m] = A[r * SC + cc + ip_tap(m)];
                v = ip_combine<V, K>(s);
            }
            Z[c] = v;
        }
        __syncthreads();
        V y;
        {
            const int cr = c1 - r0;
            if (!oy) {
                y = Z[cr * T2 + tzi];
            } else {
                V s[K];
#pragma unroll
                for (int m = 0; m < K; m++) s[m] = Z[(cr + ip_tap(m)) * T2 + tzi];
                y = ip_combine<V, K>(s);
            }
        }
#pragma unroll
        for (int m = 0; m < K - 1; m++) ring[m] = ring[m + 1];
        ring[K - 1] = y;
    }
    opass(p0e + 1);
}

}  // namespace gpaw
