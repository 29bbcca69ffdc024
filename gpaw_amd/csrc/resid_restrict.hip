// Residual of a 7-point operator and its K = 2 restriction in one launch,
// for the descent of the multigrid V-cycle below level 0 (gfx950).
//
// gpaw/poisson.py:257-261 forms r = L phi - rho on level l, restricts r onto
// rho_{l+1} and zeroes phi_{l+1}: as separate launches, r is written to
// memory by one kernel and read back by the next, and each launch pays its
// own start and drain (on the small levels most of their time).  Here a
// workgroup owns a BC^3 box of coarse outputs.  It loads phi over the
// (2 BC + 1)^3 fine cells those outputs restrict from, plus the stencil's
// R-cell halo, into LDS (zero outside a zero-boundary axis, wrapped on a
// periodic one), forms the residual of those cells in LDS and restricts it
// in three separable passes, z, then y, then x, each rounded to the working
// precision.  The residual never reaches memory.
//
// Arithmetic, so that the bits equal the two launches it replaces:
//   residual   the star stencil kernels' ST_RESID expression: x = c_0 f,
//              then the taps z+, z-, y+, y-, x+, x- (at distances 1 .. R;
//              R = 1 is instantiated) (c/bmgs/stencils.c via st_epilogue),
//              r = x - rho;
//   restrict   per pass b = 0.5 * (a_0 + 0.5 (a_+1 + a_-1)) (transfer_common.h,
//              c/bmgs/restrict.c:18-63 with K = 2), output i centred on input
//              cell 1 + 2 i - lo.
// A residual cell outside a zero-boundary axis is 0, as the restriction's
// ghost read of r gives; on a periodic axis a cell is evaluated at its
// unwrapped position from wrapped phi and rho, which is the residual of the
// wrapped cell.  Real data only (the Poisson hierarchy).
#include <cstdlib>

#include "kernels.h"
#include "stencil_tile.h"
#include "transfer_common.h"

namespace gpaw {

TrArgs make_trargs(const Transfer& tr);

namespace {

constexpr int RR_NT = 256;

// index j of an axis of extent n: wrapped (any number of periods) or
// flagged outside a zero-boundary axis
__device__ __forceinline__ int rr_wrap(int j, int n, int per, bool& out)
{
    if (j >= 0 && j < n) return j;
    if (!per) {
        out = true;
        return 0;
    }
    j %= n;
    return j < 0 ? j + n : j;
}

// one restriction pass of K = 2: 0.5 * (centre + 0.5 * (plus + minus))
template <class V>
__device__ __forceinline__ V rr_pass(V c, V p, V m)
{
    typedef typename VT<V>::R RT;
    return vmul(vmac(c, vadd(p, m), (RT)0.5), (RT)0.5);
}

template <class V, int R, int BC>
__global__ __launch_bounds__(RR_NT) void k_resid_restrict(const V* __restrict__ phi,
                                                          const V* __restrict__ rho,
                                                          V* __restrict__ rho_c,
                                                          V* __restrict__ phi_c, Geom g,
                                                          TileCoefs sc, TrArgs t, int nb1, int nb2)
{
    constexpr int F = 2 * BC + 1;   // residual cells per axis
    constexpr int P = F + 2 * R;    // phi cells per axis
    constexpr int NR = (F * F * F + RR_NT - 1) / RR_NT;
    constexpr int NP = (P * P * P + RR_NT - 1) / RR_NT;
    static_assert(NP <= 32, "phi cells per thread fit the outside mask");
    __shared__ V sp[P * P * P];
    __shared__ V sr[F * F * F];
    __shared__ V sz[F * F * BC];
    __shared__ V sy[F * BC * BC];
    typedef typename VT<V>::R RT;
    const int b = blockIdx.x;
    const int bz = b % nb2, by = (b / nb2) % nb1, bx = b / (nb2 * nb1);
    const int k0 = bx * BC, k1 = by * BC, k2 = bz * BC;   // first coarse output
    // first residual cell: the first output's centre 1 + 2k - lo, minus one
    const int f0 = 2 * k0 - t.lo[0], f1 = 2 * k1 - t.lo[1], f2 = 2 * k2 - t.lo[2];
    const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
    const i64 plane = (i64)n1 * n2;
    // every load is issued before the first is consumed: phi of this thread's
    // cells into registers, then rho, then the LDS stores
    V pv[NP];
    unsigned pout = 0;   // bit p: phi cell threadIdx.x + p * RR_NT is outside
#pragma unroll
    for (int p = 0; p < NP; p++) {
        const int i = threadIdx.x + p * RR_NT;
        pv[p] = vzero<V>();
        if (i < P * P * P) {
            const int x = i / (P * P), y = (i / P) % P, z = i % P;
            bool o = false;
            const int j0 = rr_wrap(f0 - R + x, n0, g.per[0], o);
            const int j1 = rr_wrap(f1 - R + y, n1, g.per[1], o);
            const int j2 = rr_wrap(f2 - R + z, n2, g.per[2], o);
            pv[p] = phi[o ? 0 : (i64)j0 * plane + (i64)j1 * n2 + j2];
            if (o) pout |= 1u << p;
        }
    }
    V rv[NR];
    unsigned outside = 0;   // bit p: residual cell threadIdx.x + p * RR_NT is outside
#pragma unroll
    for (int p = 0; p < NR; p++) {
        const int i = threadIdx.x + p * RR_NT;
        rv[p] = vzero<V>();
        if (i < F * F * F) {
            const int x = i / (F * F), y = (i / F) % F, z = i % F;
            bool o = false;
            const int j0 = rr_wrap(f0 + x, n0, g.per[0], o);
            const int j1 = rr_wrap(f1 + y, n1, g.per[1], o);
            const int j2 = rr_wrap(f2 + z, n2, g.per[2], o);
            const V v = rho[o ? 0 : (i64)j0 * plane + (i64)j1 * n2 + j2];
            rv[p] = v;
            if (o) outside |= 1u << p;
        }
    }
#pragma unroll
    for (int p = 0; p < NP; p++) {
        const int i = threadIdx.x + p * RR_NT;
        if (i < P * P * P) sp[i] = ((pout >> p) & 1) ? vzero<V>() : pv[p];
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < NR; p++) {
        const int i = threadIdx.x + p * RR_NT;
        if (i < F * F * F) {
            const int x = i / (F * F), y = (i / F) % F, z = i % F;
            const V* c = sp + ((x + R) * P + y + R) * P + z + R;
            V s = vzero<V>();
            s = vmac(s, c[0], (RT)sc.c[0]);
#pragma unroll
            for (int k = 1; k <= R; k++) s = vmac(s, c[k], (RT)sc.c[k]);
#pragma unroll
            for (int k = 1; k <= R; k++) s = vmac(s, c[-k], (RT)sc.c[R + k]);
#pragma unroll
            for (int k = 1; k <= R; k++) s = vmac(s, c[k * P], (RT)sc.c[2 * R + k]);
#pragma unroll
            for (int k = 1; k <= R; k++) s = vmac(s, c[-k * P], (RT)sc.c[3 * R + k]);
#pragma unroll
            for (int k = 1; k <= R; k++) s = vmac(s, c[k * P * P], (RT)sc.c[4 * R + k]);
#pragma unroll
            for (int k = 1; k <= R; k++) s = vmac(s, c[-k * P * P], (RT)sc.c[5 * R + k]);
            sr[i] = ((outside >> p) & 1) ? vzero<V>() : st_epilogue<ST_RESID>(s, c[0], rv[p], (RT)0,
                                                                             (RT)sc.c[0]);
        }
    }
    __syncthreads();
    // z pass: (fine x, fine y, coarse z)
    for (int i = threadIdx.x; i < F * F * BC; i += RR_NT) {
        const int xy = i / BC, kz = i % BC;
        const V* s = sr + xy * F + 2 * kz + 1;
        sz[i] = rr_pass(s[0], s[1], s[-1]);
    }
    __syncthreads();
    // y pass: (fine x, coarse y, coarse z)
    for (int i = threadIdx.x; i < F * BC * BC; i += RR_NT) {
        const int x = i / (BC * BC), ky = (i / BC) % BC, kz = i % BC;
        const V* s = sz + (x * F + 2 * ky + 1) * BC + kz;
        sy[i] = rr_pass(s[0], s[BC], s[-BC]);
    }
    __syncthreads();
    // x pass and the stores
    if (threadIdx.x < BC * BC * BC) {
        const int kx = threadIdx.x / (BC * BC), ky = (threadIdx.x / BC) % BC, kz = threadIdx.x % BC;
        const V* s = sy + ((2 * kx + 1) * BC + ky) * BC + kz;
        const V v = rr_pass(s[0], s[BC * BC], s[-BC * BC]);
        const int o0 = k0 + kx, o1 = k1 + ky, o2 = k2 + kz;
        if (o0 < t.nout[0] && o1 < t.nout[1] && o2 < t.nout[2]) {
            const i64 q = ((i64)o0 * t.nout[1] + o1) * t.nout[2] + o2;
            rho_c[q] = v;
            if (phi_c) phi_c[q] = vzero<V>();
        }
    }
}

// coarse outputs per block edge: 4, or 2 where 4^3 boxes would leave most of
// the chip idle (GPAW_AMD_RR_BOX=2 or 4 forces one)
int rr_box(const TrArgs& t)
{
    static const int forced = [] {
        const char* e = getenv("GPAW_AMD_RR_BOX");
        return e ? atoi(e) : 0;
    }();
    if (forced == 2 || forced == 4) return forced;
    const long nb4 = (long)((t.nout[0] + 3) / 4) * ((t.nout[1] + 3) / 4) * ((t.nout[2] + 3) / 4);
    return nb4 >= 128 ? 4 : 2;
}

template <class V, int R>
hipError_t rr_launch(const void* phi, const void* rho, void* rho_c, void* phi_c, const Geom& g,
                     const TileCoefs& sc, const TrArgs& t, hipStream_t st)
{
    const int bc = rr_box(t);
    const int nb0 = (t.nout[0] + bc - 1) / bc, nb1 = (t.nout[1] + bc - 1) / bc,
              nb2 = (t.nout[2] + bc - 1) / bc;
    const dim3 grid((unsigned)(nb0 * nb1 * nb2));
    if (bc == 4)
        launch_kernel(k_resid_restrict<V, R, 4>, grid, dim3(RR_NT), 0, st, (const V*)phi,
                      (const V*)rho, (V*)rho_c, (V*)phi_c, g, sc, t, nb1, nb2);
    else
        launch_kernel(k_resid_restrict<V, R, 2>, grid, dim3(RR_NT), 0, st, (const V*)phi,
                      (const V*)rho, (V*)rho_c, (V*)phi_c, g, sc, t, nb1, nb2);
    return launch_error();
}

}  // namespace

bool resid_restrict_ok(const Stencil& s, const Geom& g, const Transfer& tr)
{
    // 7-point operators (the levels below 0 of the Poisson hierarchy): on
    // wider stars the halo of R + 1 cells around each box makes the block
    // loads cost more than the second launch they save
    if (s.star_r != 1) return false;
    if (tr.K != 2 || tr.interpolate) return false;
    for (int a = 0; a < 3; a++) {
        if (tr.gin.n[a] != g.n[a] || tr.gin.per[a] != g.per[a]) return false;
        if (g.n[a] < 1 || tr.nout[a] < 1) return false;
    }
    return true;
}

hipError_t launch_resid_restrict(const Stencil& s, int dtype, const void* phi, const void* rho,
                                 void* rho_c, void* phi_c, const Geom& g, const Transfer& tr,
                                 hipStream_t st)
{
    if (!resid_restrict_ok(s, g, tr)) return hipErrorInvalidValue;
    TileCoefs sc{};
    for (int c = 0; c < 6 * s.star_r + 1; c++) sc.c[c] = s.coefs[c];
    const TrArgs t = make_trargs(tr);
    if (dtype == DT_F64) return rr_launch<double, 1>(phi, rho, rho_c, phi_c, g, sc, t, st);
    if (dtype == DT_F32) return rr_launch<float, 1>(phi, rho, rho_c, phi_c, g, sc, t, st);
    return hipErrorInvalidValue;
}

}  // namespace gpaw
