// Coarse-grid correction and the first post-smoothing sweep of level 0 in one
// launch (gfx950).
//
// gpaw/poisson.py:263-270 interpolates the coarser level's result onto the
// level (K = 2), subtracts it from phi and then relaxes.  As separate
// launches the correction reads and writes the whole of phi once more and
// pays its own launch; here the sweep applies it to the values it loads.
// The kernel is the box stencil kernel's ST_JACOBI sweep (stencil_box.h: a
// block of 32 x 8 threads owns a 32 x 8 tile of Q planes, loads its column
// over Q + 2R planes, the in-plane halo of its Q planes and rho, one LDS
// round) with two additions:
//   * before the barrier, the block also loads the coarse values its cells
//     interpolate from (at most CN(Q + 2R) x CN(8 + 2R) x CN(32 + 2R)) and
//     stages them in LDS, then a first barrier;
//   * every loaded phi value inside the grid becomes phi - I e, with
//     interp2_combine's arithmetic (transfer_common.h: the z, y, x passes of
//     c/bmgs/interpolate.c:19-66 with K = 2), the value the separate
//     interpolate-and-subtract launch stores there.
// The sweep's arithmetic is the box kernel's, so the output equals the two
// launches bit for bit.  Ghost cells of a zero-boundary axis stay zero;
// cells past a periodic axis are corrected at their unwrapped position,
// which on a periodic axis of twice the coarse extent is the correction of
// the wrapped cell.  Real data only (the Poisson hierarchy), one band.
#include <type_traits>

#include "stencil_box.h"
#include "transfer_common.h"

namespace gpaw {

TrArgs make_trargs(const Transfer& tr);

namespace {

// coarse cells behind L consecutive fine cells of one axis (either parity)
constexpr int jc_cn(int L) { return (L - 1) / 2 + 2; }

// coarse index j on an axis of extent n: wrapped or flagged outside a
// zero-boundary axis
__device__ __forceinline__ int jc_wrap(int j, int n, int per, bool& out)
{
    if (j >= 0 && j < n) return j;
    if (!per) {
        out = true;
        return 0;
    }
    j %= n;
    return j < 0 ? j + n : j;
}

template <class V, int R, int Q>
__global__ __launch_bounds__(BOX_NT) void k_jacobi_corr(const V* __restrict__ a, V* __restrict__ o,
                                                        const V* __restrict__ sa,
                                                        const V* __restrict__ corr, Geom g,
                                                        TileCoefs sc, double w, TileShape sh,
                                                        TrArgs ct)
{
    typedef typename VT<V>::R RT;
    typedef BoxGeom<V, R, Q> BG;
    constexpr int TZ = BG::TZ, TY = BG::TY, LW = BG::LW, LH = BG::LH, NHALO = BG::NHALO,
                  NH = BG::NH;
    constexpr int CX = jc_cn(Q + 2 * R), CY = jc_cn(TY + 2 * R), CZ = jc_cn(TZ + 2 * R);
    constexpr int NCS = CX * CY * CZ;
    constexpr int NC = (NCS + BOX_NT - 1) / BOX_NT;
    static_assert(NC <= 32, "coarse values per thread fit the outside mask");
    __shared__ V lds[BG::LDS];
    __shared__ V cs[NCS];

    const int b = blockIdx.x;
    const int t = xcd_tile(b, sh.nblocks);
    const int tzi = t % sh.ntz;
    int rest = t / sh.ntz;
    const int tyb = rest % sh.nty;
    const int ch = rest / sh.nty;

    const int tx = threadIdx.x, tyi = threadIdx.y;
    const int tid = tx + TZ * tyi;
    const int n0 = g.n[0];
    const i64 plane = (i64)g.n[1] * g.n[2];
    const int z0 = tzi * TZ, y0 = tyb * TY;
    const int i0s = ch * sh.q;
    const int i1 = y0 + tyi, i2 = z0 + tx;
    const bool inr = (i1 < g.n[1]) && (i2 < g.n[2]);
    const PlaneRef qr = plane_ref(i1, i2, g);

    int hoff[NH];
    PlaneRef hr[NH];
#pragma unroll
    for (int k = 0; k < NH; k++) {
        const int h = tid + BOX_NT * k;
        int lr, lc;
        if (h < 2 * R * TZ) {
            const int yr = h / TZ;
            lc = R + h % TZ;
            lr = yr < R ? yr : TY + yr;
        } else {
            const int hh = h - 2 * R * TZ;
            const int zc = hh % (2 * R);
            lr = R + hh / (2 * R);
            lc = zc < R ? zc : TZ + zc;
        }
        hoff[k] = (h < NHALO) ? lr * LW + lc : -1;
        hr[k] = (h < NHALO) ? plane_ref(y0 + lr - R, z0 + lc - R, g) : 0;
    }

    // the coarse values of the block's cells: unwrapped coarse indices from cm
    const int cm0 = ((i0s - R + ct.skip[0][0]) >> 1) - ct.lo[0];
    const int cm1 = ((y0 - R + ct.skip[1][0]) >> 1) - ct.lo[1];
    const int cm2 = ((z0 - R + ct.skip[2][0]) >> 1) - ct.lo[2];
    V cv[NC];
    unsigned cmask = 0;   // bit p: coarse cell tid + p * BOX_NT is a zero ghost
#pragma unroll
    for (int p = 0; p < NC; p++) {
        const int i = tid + p * BOX_NT;
        cv[p] = vzero<V>();
        if (i < NCS) {
            bool out = false;
            const int c0 = jc_wrap(cm0 + i / (CY * CZ), ct.gin.n[0], ct.gin.per[0], out);
            const int c1 = jc_wrap(cm1 + (i / CZ) % CY, ct.gin.n[1], ct.gin.per[1], out);
            const int c2 = jc_wrap(cm2 + i % CZ, ct.gin.n[2], ct.gin.per[2], out);
            cv[p] = corr[out ? 0 : ((i64)c0 * ct.gin.n[1] + c1) * ct.gin.n[2] + c2];
            if (out) cmask |= 1u << p;
        }
    }

    // the box kernel's loads: column over Q + 2R planes, halo and rho of the
    // Q planes
    V q[Q + 2 * R], hq[Q][NH], sn[Q];
    int qf[Q + 2 * R];
#pragma unroll
    for (int p = 0; p < Q + 2 * R; p++) {
        int j = i0s - R + p;
        qf[p] = wrap_index(j, n0, g.per[0]);
        const i64 base = qf[p] == 3 ? 0 : (i64)j * plane;
        q[p] = a[base + pref_off(qr)];
    }
#pragma unroll
    for (int p = 0; p < Q; p++) {
        const i64 base = qf[R + p] == 3 ? 0 : (i64)(i0s + p - (qf[R + p] == 1 ? n0 : 0)) * plane;
#pragma unroll
        for (int k = 0; k < NH; k++) hq[p][k] = a[base + pref_off(hr[k])];
        const bool ok = inr && i0s + p < n0;
        sn[p] = sa[ok ? (i64)(i0s + p) * plane + pref_off(qr) : 0];
    }
#pragma unroll
    for (int p = 0; p < NC; p++)
        if (tid + p * BOX_NT < NCS) cs[tid + p * BOX_NT] = ((cmask >> p) & 1) ? vzero<V>() : cv[p];
    __syncthreads();

    // phi - I e, pass by pass as interp2_combine: the z and y passes of a
    // cell depend on its (i1, i2) and one coarse x index, so each thread
    // forms them once per coarse x index of its column and of its halo cell
    // and combines along x per plane.  The x parity of plane p of the block
    // is (sx + p) & 1 with sx uniform: one instantiation per value keeps the
    // register indices constant.
    const RT h = (RT)0.5;
    // y-z value at coarse x index c of a cell with in-plane position (f1, f2)
    auto yz_at = [&](int c, int l1, int l2, bool py, bool pz) -> V {
        const V* r0 = cs + (c * CY + l1) * CZ + l2;
        const V* r1 = r0 + (py ? CZ : 0);
        const int dz = pz ? 1 : 0;
        const V z0 = pz ? vmul(vadd(r0[0], r0[dz]), h) : r0[0];
        const V z1 = pz ? vmul(vadd(r1[0], r1[dz]), h) : r1[0];
        return py ? vmul(vadd(z0, z1), h) : z0;
    };
    auto corrections = [&](auto SXC) {
        constexpr int SX = decltype(SXC)::value;
        // the column: fine planes i0s - R + p, coarse x index (SX + p) >> 1
        {
            const int p1 = i1 + ct.skip[1][0], p2 = i2 + ct.skip[2][0];
            const int l1 = (p1 >> 1) - ct.lo[1] - cm1, l2 = (p2 >> 1) - ct.lo[2] - cm2;
            V yz[CX];
#pragma unroll
            for (int c = 0; c < CX; c++) yz[c] = yz_at(c, l1, l2, p1 & 1, p2 & 1);
            const bool col_out = ((qr & 3) == 3) || (((qr >> 2) & 3) == 3);
#pragma unroll
            for (int p = 0; p < Q + 2 * R; p++) {
                const int c = (SX + p) >> 1;
                const V e = ((SX + p) & 1) ? vmul(vadd(yz[c], yz[c + 1]), h) : yz[c];
                q[p] = ghost_fix(q[p], qf[p], qr, Phases{});
                if (qf[p] != 3 && !col_out) q[p] = vsub(q[p], e);
            }
        }
        // the halo cells: planes i0s + p (column planes R + p)
#pragma unroll
        for (int k = 0; k < NH; k++) {
            if (hoff[k] < 0) continue;
            const int f1 = y0 + hoff[k] / LW - R, f2 = z0 + hoff[k] % LW - R;
            const int p1 = f1 + ct.skip[1][0], p2 = f2 + ct.skip[2][0];
            const int l1 = (p1 >> 1) - ct.lo[1] - cm1, l2 = (p2 >> 1) - ct.lo[2] - cm2;
            constexpr int C0 = (SX + R) >> 1, C1 = (SX + R + Q) >> 1;
            V yz[C1 - C0 + 1];
#pragma unroll
            for (int c = C0; c <= C1; c++) yz[c - C0] = yz_at(c, l1, l2, p1 & 1, p2 & 1);
            const bool hout = (hr[k] & 3) == 3 || ((hr[k] >> 2) & 3) == 3;
#pragma unroll
            for (int p = 0; p < Q; p++) {
                const int c = ((SX + R + p) >> 1) - C0;
                const V e = ((SX + R + p) & 1) ? vmul(vadd(yz[c], yz[c + 1]), h) : yz[c];
                V v = ghost_fix(hq[p][k], qf[R + p], hr[k], Phases{});
                if (qf[R + p] != 3 && !hout) v = vsub(v, e);
                hq[p][k] = v;
            }
        }
    };
    if ((i0s - R + ct.skip[0][0]) & 1)
        corrections(std::integral_constant<int, 1>());
    else
        corrections(std::integral_constant<int, 0>());
#pragma unroll
    for (int p = 0; p < Q; p++) {
        V* B = lds + p * (LH * LW);
        B[(R + tyi) * LW + R + tx] = q[R + p];
#pragma unroll
        for (int k = 0; k < NH; k++)
            if (hoff[k] >= 0) B[hoff[k]] = hq[p][k];
    }
    __syncthreads();

#pragma unroll
    for (int p = 0; p < Q; p++) {
        const V* row = lds + p * (LH * LW) + (R + tyi) * LW + R + tx;
        const V f = q[R + p];
        V x = vzero<V>();
#pragma unroll
        for (int k = 1; k <= R; k++) x = vmac(x, row[k], (RT)sc.c[k]);
#pragma unroll
        for (int k = 1; k <= R; k++) x = vmac(x, row[-k], (RT)sc.c[R + k]);
#pragma unroll
        for (int k = 1; k <= R; k++) x = vmac(x, row[k * LW], (RT)sc.c[2 * R + k]);
#pragma unroll
        for (int k = 1; k <= R; k++) x = vmac(x, row[-k * LW], (RT)sc.c[3 * R + k]);
#pragma unroll
        for (int k = 1; k <= R; k++) x = vmac(x, q[R + p + k], (RT)sc.c[4 * R + k]);
#pragma unroll
        for (int k = 1; k <= R; k++) x = vmac(x, q[R + p - k], (RT)sc.c[5 * R + k]);
        const V r = st_epilogue<ST_JACOBI>(x, f, sn[p], (RT)w, (RT)sc.c[0]);
        if (inr && i0s + p < n0) o[(i64)(i0s + p) * plane + pref_off(qr)] = r;
    }
}

// planes per block
#ifndef GPAW_JC_Q_F32
#define GPAW_JC_Q_F32 8
#endif
#ifndef GPAW_JC_Q_F64
#define GPAW_JC_Q_F64 4
#endif
template <class V>
constexpr int jc_q()
{
    return sizeof(V) == 4 ? GPAW_JC_Q_F32 : GPAW_JC_Q_F64;
}

template <class V, int R>
hipError_t jc_launch(const Stencil& s, const void* in, void* out, const void* rho,
                     const void* corr, const Geom& g, double w, const TileShape& sh,
                     const TrArgs& ct, hipStream_t st)
{
    TileCoefs sc;
    for (int c = 0; c < 6 * R + 1; c++) sc.c[c] = s.coefs[c];
    launch_kernel((k_jacobi_corr<V, R, jc_q<V>()>), dim3(sh.nblocks), dim3(BOX_TZ, BOX_NTY), 0, st,
                  (const V*)in, (V*)out, (const V*)rho, (const V*)corr, g, sc, w, sh, ct);
    return launch_error();
}

}  // namespace

bool jacobi_corr_ok(const Stencil& s, const Geom& g, const Transfer& itp)
{
    if (s.star_r < 1 || s.star_r > 3) return false;
    if (itp.K != 2 || !itp.interpolate) return false;
    for (int a = 0; a < 3; a++) {
        if (itp.nout[a] != g.n[a]) return false;
        // a periodic fine axis wraps where its coarse axis does
        if (g.per[a] && (!itp.gin.per[a] || g.n[a] != 2 * itp.gin.n[a])) return false;
    }
    TileShape sh;
    return box_geometry(g, s.star_r, 1, jc_q<float>(), sh) &&
           box_geometry(g, s.star_r, 1, jc_q<double>(), sh);
}

hipError_t launch_jacobi_corr(const Stencil& s, int dtype, const void* in, void* out,
                              const void* rho, const void* corr, const Geom& g, double w,
                              const Transfer& itp, hipStream_t st)
{
    if (!jacobi_corr_ok(s, g, itp)) return hipErrorInvalidValue;
    TileShape sh;
    box_geometry(g, s.star_r, 1, dtype == DT_F32 ? jc_q<float>() : jc_q<double>(), sh);
    const TrArgs ct = make_trargs(itp);
#define GPAW_JC(T)                                                                   \
    switch (s.star_r) {                                                              \
    case 1: return jc_launch<T, 1>(s, in, out, rho, corr, g, w, sh, ct, st);         \
    case 2: return jc_launch<T, 2>(s, in, out, rho, corr, g, w, sh, ct, st);         \
    default: return jc_launch<T, 3>(s, in, out, rho, corr, g, w, sh, ct, st);        \
    }
    if (dtype == DT_F64) GPAW_JC(double)
    if (dtype == DT_F32) GPAW_JC(float)
#undef GPAW_JC
    return hipErrorInvalidValue;
}

}  // namespace gpaw
